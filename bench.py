#!/usr/bin/env python3
"""Benchmark of the FIPI time step on one MI355X.

Workload (the reference's real-scale configuration): 256^3 periodic box,
dx = 1, Cahn-Hilliard (eps = 1, M = 1, kappa = 1) coupled to Stokes
(eta = 1), semi-implicit spectral step dt = 0.01; phi0 = U(-0.05, 0.05)
from a seed; 100,000 rigid spheres of radius 2 dx placed by seeded random
sequential addition (gap 0.5 dx), contact angles U(60, 120) deg.

One step = one pass of the reference time loop body (src/main.cpp:75-222):
CH SBDF2 update, Stokes solve, two Heun particle stages.  Inputs are
resident in HBM before the timed region.  Prints one JSON line.

The line also carries the batched-ensemble measurement ("ensemble"): 512
independent boxes with edges drawn from {32, 48, 64, 80, 96}, mean
composition U(-0.3, 0.3) + U(-0.05, 0.05) noise, U(0, 2000) spheres of
radius 2 dx per box (capped where random sequential addition jams),
FP32-mixed, advanced together (one graph replay per step), against the
same boxes stepped one at a time through the single-box path.

And the calibration adjoint ("adjoint"): 128^3 box, 20,000 spheres of
radius 2 dx held fixed, 200 steps; target from a seeded forward run with
M = 1.3, kappa = 0.8 and contact angles U(70, 110) deg; loss = sum of
squared phi mismatch at the last step; gradients in M, kappa and every
contact angle from the hand-written adjoint (checkpoints every 20 steps),
FP32-mixed and FP64, checked against FP64 central differences (h = 1e-4).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "steps/s at 256^3+100k particles (FP64 vs FP32-mixed @ stated err), HBM GB/s/peak"
HBM_PEAK_GBS = 8000.0


def workload(n, n_particles, precision, seed=1):
    import fipi_amd
    p = fipi_amd.config_params(n, n_particles, 2.0, 1000, precision=precision)
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, seed)
    pos = fipi_amd.init_rsa(n_particles, 2.0, 0.5, (p.lx, p.ly, p.lz), seed + 1)
    if pos.shape[0] != n_particles:
        raise SystemExit(f"RSA placed only {pos.shape[0]} of {n_particles} spheres")
    theta = fipi_amd.init_uniform(n_particles, 60.0, 120.0, seed + 2)
    cos_theta = np.cos(np.deg2rad(theta))
    return p, phi0, pos, cos_theta


EDGE_SET = (32, 48, 64, 80, 96)


def ensemble_workload(nboxes, seed=11, max_attempts=2000):
    """Seeded boxes of the ensemble workload: (edge, spheres placed,
    phi0, positions) per box, the requested counts and how many boxes
    jammed below their request."""
    import fipi_amd
    rng = np.random.default_rng(seed)
    edges = rng.choice(EDGE_SET, size=nboxes)
    means = rng.uniform(-0.3, 0.3, size=nboxes)
    want = rng.integers(0, 2001, size=nboxes)
    boxes, capped = [], 0
    for b in range(nboxes):
        n = int(edges[b])
        phi = fipi_amd.init_phi((n, n, n), float(means[b]), 0.05, seed * 100003 + b)
        pos = fipi_amd.init_rsa(int(want[b]), 2.0, 0.5, (n, n, n), seed * 200003 + b, max_attempts)
        capped += int(pos.shape[0] < want[b])
        boxes.append((n, int(pos.shape[0]), phi, pos))
    return boxes, want, capped


def ensemble_run(nboxes, steps, loop_sample, loop_steps):
    """Aggregate grid-point-steps/s of the batched ensemble and of looping
    the single-box path over a sample of the same boxes."""
    import torch
    import fipi_amd
    boxes, want, capped = ensemble_workload(nboxes)
    common = fipi_amd.config_params(32, 0, 2.0, steps, precision=fipi_amd.FP32_MIXED).replace(switch_particle=1)
    sum_g = sum(n ** 3 for n, _, _, _ in boxes)
    ens = fipi_amd.Ensemble(common, [b[0] for b in boxes], [b[1] for b in boxes])
    for i, (n, k, phi, pos) in enumerate(boxes):
        ens.set_box(i, phi, pos if k else None)
    t0 = time.perf_counter()
    ens.tune(3)
    tune_s = time.perf_counter() - t0
    ens.step(2)  # first-order start step, then graph capture + first replay
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step_async(steps)
    ens.sync()
    wall = time.perf_counter() - t0
    m0 = [ens.diagnostics(b)["mean_phi"] for b in range(min(nboxes, 8))]
    tuning = {str(n): ens.tuning(n)
              for n in EDGE_SET}
    ens.close()
    del ens
    torch.cuda.synchronize()
    # the same boxes one at a time through the single-box path
    loop_g, loop_t = 0, 0.0
    for n, k, phi, pos in boxes[:loop_sample]:
        sim = fipi_amd.Simulation(fipi_amd.config_params(n, k, 2.0, steps, precision=fipi_amd.FP32_MIXED))
        sim.set_state(phi, pos if k else None)
        sim.step(2)
        t0 = time.perf_counter()
        sim.step_async(loop_steps)
        sim.sync()
        loop_t += time.perf_counter() - t0
        loop_g += n ** 3 * loop_steps
        sim.close()
    rate = sum_g * steps / wall
    loop_rate = loop_g / loop_t
    hist = {str(n): int(sum(1 for b in boxes if b[0] == n)) for n in EDGE_SET}
    return {
        "workload": f"{nboxes} boxes, edges from {list(EDGE_SET)}, U(0,2000) spheres r=2dx (RSA-capped), "
                    f"FP32-mixed, dt=0.01",
        "boxes_per_edge": hist,
        "grid_points": sum_g,
        "particles": int(sum(b[1] for b in boxes)),
        "particles_requested": int(want.sum()),
        "boxes_capped_by_rsa": capped,
        "steps": steps,
        "ms_per_step": round(wall * 1e3 / steps, 4),
        "grid_point_steps_per_s": rate,
        "looped_single_box_grid_point_steps_per_s": loop_rate,
        "looped_sample": f"first {loop_sample} boxes, {loop_steps} steps each after 2 untimed",
        "speedup_vs_looped": round(rate / loop_rate, 2),
        "tuning": tuning,
        "tune_seconds": round(tune_s, 2),
        "mean_phi_first_boxes": m0,
    }


def adjoint_run(n=128, npart=20000, steps=200, every=20, wetting=0.5, seed=21):
    """Gradient timing and checks of the calibration adjoint."""
    import fipi_amd
    from fipi_amd import FP32_MIXED, FP64
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, seed)
    pos = fipi_amd.init_rsa(npart, 2.0, 0.0, (n, n, n), seed + 1)
    th_true = np.deg2rad(fipi_amd.init_uniform(pos.shape[0], 70.0, 110.0, seed + 2))
    npart = pos.shape[0]
    th0 = np.full(npart, np.pi / 2)

    def make(prec):
        p = fipi_amd.config_params(n, npart, 2.0, steps, precision=prec)
        return fipi_amd.Adjoint(p, phi0, pos, wetting, steps, every)

    a32 = make(FP32_MIXED)
    _, target = a32.forward(1.3, 0.8, th_true, want_loss=False)
    a32.set_target(target)
    t0 = time.perf_counter()
    a32.forward(1.0, 1.0, th0)
    fwd32 = time.perf_counter() - t0
    a32.gradient(1.0, 1.0, th0)  # first use of the backward kernels
    t0 = time.perf_counter()
    l32, m32, k32, g32 = a32.gradient(1.0, 1.0, th0)
    grad32 = time.perf_counter() - t0
    a32.close()
    a64 = make(FP64)
    a64.set_target(target)
    t0 = time.perf_counter()
    l64, m64, k64, g64 = a64.gradient(1.0, 1.0, th0)
    grad64 = time.perf_counter() - t0
    h = 1e-4

    def loss(mm, kk, tt):
        return a64.forward(mm, kk, tt)[0]
    fd_m = (loss(1.0 + h, 1.0, th0) - loss(1.0 - h, 1.0, th0)) / (2 * h)
    fd_k = (loss(1.0, 1.0 + h, th0) - loss(1.0, 1.0 - h, th0)) / (2 * h)
    fd_t = []
    for i in np.argsort(-np.abs(g64))[:3]:
        tp, tm = th0.copy(), th0.copy()
        tp[i] += h
        tm[i] -= h
        fd_t.append((int(i), float(g64[i]), (loss(1.0, 1.0, tp) - loss(1.0, 1.0, tm)) / (2 * h)))
    a64.close()
    gs = float(np.abs(g64).max())
    return {
        "workload": f"{n}^3, {npart} spheres r=2dx fixed, {steps} steps, checkpoints every {every}, "
                    f"wetting a_w={wetting}; target M=1.3 kappa=0.8 theta~U(70,110) deg; gradient at "
                    f"M=1 kappa=1 theta=90 deg",
        "loss_fp64": l64,
        "forward_ms_fp32_mixed": round(fwd32 * 1e3, 2),
        "gradient_ms_fp32_mixed": round(grad32 * 1e3, 2),
        "gradient_ms_fp64": round(grad64 * 1e3, 2),
        "grad_fp64": {"mobility": m64, "kappa": k64, "theta_norm": float(np.linalg.norm(g64))},
        "fd_rel_err_fp64": {
            "mobility": abs(m64 - fd_m) / abs(fd_m),
            "kappa": abs(k64 - fd_k) / abs(fd_k),
            "theta_top3_vs_max_grad": max(abs(a - b) for _, a, b in fd_t) / gs,
        },
        "mixed_vs_fp64_rel": {
            "loss": abs(l32 - l64) / l64,
            "mobility": abs(m32 - m64) / abs(m64),
            "kappa": abs(k32 - k64) / abs(k64),
            "theta": float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)),
        },
    }


def step_bytes(p, n_particles):
    """Algorithmic HBM bytes of one step with the minimal dataflow
    (every array read or written once per stage): 20 real fields + 18 half
    spectra + 2 Heun stages of particle state, gathers and the spread."""
    s = 8 if p.precision == 0 else 4
    R = p.G * s
    H = p.nz * p.ny * (p.nx // 2 + 1) * 2 * s
    bp = 2 * (6 * 24 + 32 * s) + 24 * s
    return 20 * R + 18 * H + n_particles * bp


def precision_run(p, phi0, pos, cos_theta, steps, fixture_path):
    """FP32-mixed against FP64 after `steps` steps from the same state
    (outside the timed region): relative L2 of phi, relative error of the
    interface area, largest sphere displacement, mean phi; and both
    precisions against the FP64 oracle's checkpoints where a long-horizon
    fixture of this workload exists (phi at seeded points, positions)."""
    import fipi_amd
    out = {}
    for name, prec in (("fp64", fipi_amd.FP64), ("fp32_mixed", fipi_amd.FP32_MIXED)):
        sim = fipi_amd.Simulation(p.replace(precision=prec))
        sim.set_state(phi0, pos, None, cos_theta)
        sim.step(steps)
        out[name] = (sim.phi(), sim.particles()["p"], sim.diagnostics())
        sim.close()
    (a_phi, a_p, a_d), (m_phi, m_p, m_d) = out["fp64"], out["fp32_mixed"]
    L = np.array([p.lx, p.ly, p.lz])
    dp = m_p - a_p
    dp -= np.round(dp / L) * L
    res = {
        "after_steps": steps,
        "phi_rel_l2": float(np.linalg.norm(m_phi - a_phi) / np.linalg.norm(a_phi)),
        "interface_area_rel_err": float(abs(m_d["interface_area"] / a_d["interface_area"] - 1.0)),
        "max_particle_disp_err": float(np.abs(dp).max()) if dp.size else 0.0,
        "mean_phi_abs_err": float(abs(m_d["mean_phi"] - a_d["mean_phi"])),
    }
    if fixture_path and os.path.exists(fixture_path):
        fx = np.load(fixture_path)
        if int(fx["particles"]) != pos.shape[0]:
            return res  # a fixture of another workload
        cks = [int(t) for t in fx["ck_steps"]]
        pss = [int(t) for t in fx["pos_steps"]]
        if steps in cks:
            k = cks.index(steps)
            idx, rphi = fx["sample_idx"], fx["ck_phi"][k]
            vo = {"fixture": os.path.relpath(fixture_path, ROOT), "step": steps}
            for name, (phi, pp, _) in out.items():
                ph = phi.reshape(-1)[idx]
                vo[f"{name}_phi_rel_l2_at_samples"] = float(np.linalg.norm(ph - rphi) / np.linalg.norm(rphi))
                if steps in pss:
                    rp = fx["pos"][pss.index(steps)]
                    d = pp[fx["pos_idx"]] - rp
                    d -= np.round(d / L) * L
                    vo[f"{name}_positions_rel_l2"] = float(np.linalg.norm(d) / np.linalg.norm(rp))
                    vo[f"{name}_max_particle_disp"] = float(np.abs(d).max())
            res["vs_oracle"] = vo
    return res


def host_cpu():
    """CPU model name and the host cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def oracle_rate(p, phi0, pos, cos_theta, steps, threads):
    """steps/s of the oracle after one untimed first-order step."""
    import oracle
    sim = oracle.Sim(p, phi0, pos, None, cos_theta, n_threads=threads)
    sim.step(1)
    t0 = time.perf_counter()
    sim.step(steps)
    dt = time.perf_counter() - t0
    del sim
    return steps / dt, dt


def cpu_baseline(p, phi0, pos, cos_theta, steps):
    """The oracle (C restatement of the reference step, 25 complex 3-D FFTs
    per step like the reference) timed on this host's cores, with all of
    them (at most 16) and with the reference's 4 FFTW threads."""
    model, ncpu = host_cpu()
    threads = min(ncpu, 16)
    rate, dt = oracle_rate(p, phi0, pos, cos_theta, steps, threads)
    rate4, dt4 = oracle_rate(p, phi0, pos, cos_theta, steps, 4)
    return {
        "value": rate,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} reference-structured FP64 steps (25 complex 3-D FFTs each, "
                  f"OpenMP) of the same {p.nx}^3 + {pos.shape[0]} workload after 1 warm-up step, "
                  f"{dt:.1f} s wall",
        "value_4_threads": rate4,
        "host_cpu": model,
        "host_cores_available": ncpu,
        "fft": "the oracle's Stockham FFT (FFTW is not installed)",
    }


def config1_run(steps=200, seed=7):
    """The reference's CPU-runnable case at full length: 64^3, spheres of
    radius 2.5 dx (2,000 requested, capped at the count random sequential
    addition places: they exceed random packing), FP64, 200 steps, on the
    GPU and in the oracle (all host cores), with the parity of the two after
    the last step."""
    import fipi_amd
    model, ncpu = host_cpu()
    n, requested = 64, 2000
    pos = fipi_amd.init_rsa(requested, 2.5, 0.5, (float(n),) * 3, seed + 1000)
    npart = pos.shape[0]
    p = fipi_amd.config_params(n, npart, 2.5, steps)
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, seed)
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0, pos)
    t0 = time.perf_counter()
    sim.step(steps)
    gpu_dt = time.perf_counter() - t0
    g_phi, g_pos = sim.phi(), sim.particles()["p"]
    sim.close()
    import oracle
    threads = min(ncpu, 16)
    ref = oracle.Sim(p, phi0, pos, n_threads=threads)
    t0 = time.perf_counter()
    ref.step(steps)
    cpu_dt = time.perf_counter() - t0
    r_phi, r_pos = ref.phi(), ref.particles()["p"]
    del ref
    return {
        "workload": f"{n}^3 + {npart} spheres r=2.5dx ({requested} requested, capped at the RSA count), FP64, "
                    f"{steps} steps from t = 0",
        "gpu_steps_per_s": steps / gpu_dt,
        "cpu_steps_per_s": steps / cpu_dt,
        "cpu_cores": threads,
        "gpu_vs_cpu": round(cpu_dt / gpu_dt, 1),
        "phi_rel_l2_vs_oracle": float(np.linalg.norm(g_phi - r_phi) / np.linalg.norm(r_phi)),
        "pos_rel_l2_vs_oracle": float(np.linalg.norm(g_pos - r_pos) / np.linalg.norm(r_pos)),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=980)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--particles", type=int, default=100000)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--precision-steps", type=int, default=1000,
                    help="steps of the FP32-mixed vs FP64 comparison (outside the timed region)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--single", action="store_true", help="skip the other precision")
    ap.add_argument("--ensemble-boxes", type=int, default=512)
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--ensemble-loop-sample", type=int, default=64)
    ap.add_argument("--ensemble-loop-steps", type=int, default=10)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-config1", action="store_true")
    ap.add_argument("--no-repeats", dest="repeats", action="store_false",
                    help="time one region of K steps only (default: three, median reported)")
    ap.add_argument("--no-field-only", action="store_true")
    ap.add_argument("--tune-steps", type=int, default=3,
                    help="steps per tile-width trial before the run (0: defaults)")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this size first (a kernel of known "
                         "traffic for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build measures one GPU only (one box, one MI355X); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import torch
    import fipi_amd

    if args.calibration_copy_mib > 0:
        src = torch.ones(args.calibration_copy_mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    prec = fipi_amd.FP64 if args.precision == "fp64" else fipi_amd.FP32_MIXED
    other = fipi_amd.FP32_MIXED if prec == fipi_amd.FP64 else fipi_amd.FP64
    p, phi0, pos, cos_theta = workload(args.n, args.particles, prec)

    def timed_run(precision):
        """W untimed + K timed steps from the seeded state; returns the rate,
        the final state and (for the headline precision) the kernel profile."""
        q = p.replace(precision=precision)
        sim = fipi_amd.Simulation(q)
        sim.set_state(phi0, pos, None, cos_theta)
        tuning = sim.tune(args.tune_steps) if args.tune_steps > 0 else sim.tuning()
        sim.step(args.warmup)
        stream = torch.cuda.ExternalStream(sim.stream_ptr)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        sim.sync()
        t0 = time.perf_counter()
        e0.record(stream)
        sim.step_async(args.steps)
        e1.record(stream)
        sim.sync()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # two more timed regions of K steps, continuing the run: the median
        # of the three is reported beside the first (the headline value)
        repeats = [args.steps / wall]
        for _ in range(2 if args.repeats else 0):
            sim.sync()
            r0 = time.perf_counter()
            sim.step_async(args.steps)
            sim.sync()
            repeats.append(args.steps / (time.perf_counter() - r0))
        res = {
            "tuning": tuning,
            "wall": wall,
            "repeats": repeats,
            "ev_ms": e0.elapsed_time(e1),
            "phi": sim.phi(),
            "pos": sim.particles()["p"],
            "diag": sim.diagnostics(),
        }
        if args.profile_steps > 0:
            res["stats"] = sim.profile_steps(args.profile_steps)
        sim.close()
        del sim
        torch.cuda.synchronize()
        return res

    main_run = timed_run(prec)
    other_run = timed_run(other) if not args.single else None

    def field_only_rate(precision, steps=300):
        """Steps/s of the same grid and phi0 without spheres (the field
        passes alone), after the same warm-up."""
        q = p.replace(precision=precision, n_particles=0, switch_particle=0)
        sim = fipi_amd.Simulation(q)
        sim.set_state(phi0)
        sim.step(args.warmup)
        sim.sync()
        t0 = time.perf_counter()
        sim.step_async(steps)
        sim.sync()
        dt = time.perf_counter() - t0
        sim.close()
        return steps / dt

    field_only = None
    if not args.single and not args.no_field_only and args.particles > 0:
        f64 = field_only_rate(fipi_amd.FP64)
        f32 = field_only_rate(fipi_amd.FP32_MIXED)
        field_only = {"workload": f"{args.n}^3, same phi0, no spheres, 300 steps",
                      "fp64_steps_per_s": round(f64, 3), "fp32_mixed_steps_per_s": round(f32, 3),
                      "mixed_speedup": round(f32 / f64, 3)}

    wall = main_run["wall"]
    ms_per_step = wall * 1e3 / args.steps
    value = args.steps / wall
    def summarize(stats):
        for s in stats:
            s["avg_ms"] = s["total_ms"] / max(s["launches"], 1)
            s["bytes_per_launch"] = s["bytes"] / max(s["launches"], 1)
            s["gbs"] = s["bytes_per_launch"] / (s["avg_ms"] * 1e-3) / 1e9 if s["avg_ms"] > 0 else 0.0
        stats.sort(key=lambda s: -s["total_ms"])
        return stats

    def kernel_table(stats):
        return [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in s.items()
                 if k in ("name", "launches", "avg_ms", "gbs")} for s in stats[:12]]

    stats = summarize(main_run.get("stats", []))
    # bytes the passes move per step (sum over launch sites of their algorithmic bytes)
    moved = sum(s["bytes"] for s in stats) / args.profile_steps if stats else None
    roofline = None
    if stats:
        dom = stats[0]
        traffic = None
        tr_path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
        if os.path.exists(tr_path):
            with open(tr_path) as f:
                tr = json.load(f)
            rec = tr.get(f"{args.precision}:{dom['name']}")
            traffic = rec["bytes"] if rec else None
        roofline = {
            "bound": "hbm",
            "kernel": dom["name"],
            "achieved": round(dom["gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom["gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": dom["bytes_per_launch"],
            "avg_launch_ms": round(dom["avg_ms"], 5),
        }
    bstep = step_bytes(p, args.particles)
    diag = main_run["diag"]

    precision_cmp = None
    if other_run is not None:
        f64 = main_run if prec == fipi_amd.FP64 else other_run
        mix = other_run if prec == fipi_amd.FP64 else main_run
        precision_cmp = {
            "fp64_steps_per_s": round(args.steps / f64["wall"], 3),
            "fp32_mixed_steps_per_s": round(args.steps / mix["wall"], 3),
            "mixed_speedup": round(f64["wall"] / mix["wall"], 3),
            "median_fp64_steps_per_s": round(float(np.median(f64["repeats"])), 3),
            "median_fp32_mixed_steps_per_s": round(float(np.median(mix["repeats"])), 3),
            "median_mixed_speedup": round(float(np.median(mix["repeats"]) / np.median(f64["repeats"])), 3),
        }
        if args.precision_steps > 0:
            fixture = os.path.join(ROOT, "tests", "golden", f"long_{args.n}.npz")
            precision_cmp.update(precision_run(p, phi0, pos, cos_theta, args.precision_steps, fixture))

    ens = None
    if not args.no_ensemble and args.ensemble_boxes > 0:
        ens = ensemble_run(args.ensemble_boxes, args.ensemble_steps, args.ensemble_loop_sample,
                           args.ensemble_loop_steps)

    adj = None if args.no_adjoint else adjoint_run()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p, phi0, pos, cos_theta, args.cpu_steps)
    cfg1 = None if (args.no_config1 or args.no_cpu_baseline) else config1_run()

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == fipi_amd.FP64 else "f32-mixed",
        "data": "synthetic: phi0 = U(-0.05,0.05) (seed 1), RSA spheres r=2dx gap 0.5dx (seed 2), "
                "contact angles U(60,120) deg (seed 3)",
        "config": {
            "workload": f"{args.n}^3 periodic CH+Stokes + {args.particles} spheres r=2dx, dt=0.01, "
                        f"{'FP64' if prec == fipi_amd.FP64 else 'FP32-mixed'}",
            "grid": args.n,
            "particles": args.particles,
            "precision": args.precision,
        },
        "event_ms_per_step": round(main_run["ev_ms"] / args.steps, 4),
        "repeats_steps_per_s": [round(v, 3) for v in main_run["repeats"]],
        "median_steps_per_s": round(float(np.median(main_run["repeats"])), 3),
        "precision_comparison": precision_cmp,
        "field_only": field_only,
        "grid_point_steps_per_s": value * p.G,
        "step_roofline": {
            "bytes_per_step": bstep,
            "achieved_gbs": round(bstep * value / 1e9, 1),
            "frac": round(bstep * value / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_moved_per_step": moved,
            "moved_gbs": round(moved * value / 1e9, 1) if moved else None,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "kernels": kernel_table(stats),
        "kernels_other_precision": kernel_table(summarize(other_run.get("stats", []))) if other_run else None,
        "tuning": {"headline": main_run["tuning"], "other_precision": other_run["tuning"] if other_run else None},
        "diagnostics": {k: diag[k] for k in ("t", "mean_phi", "free_energy", "interface_area")},
        "ensemble": ens,
        "adjoint": adj,
        "config1": cfg1,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
