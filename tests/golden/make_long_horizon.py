#!/usr/bin/env python3
"""Writes the long-horizon FP64 oracle fixtures (tests/golden/long_*.npz).

Two oracle runs of the same seeded workload, one with radix-4 FFT stages
(the default) and one with radix-2 stages only, advance in lockstep for
`steps` steps.  Both compute the same DFTs; they differ in round-off only.
The fixture records

* per step: the largest sphere displacement between the two runs (minimum
  image), which sphere it is, and the positions' relative L2 difference;
* every `every` steps: the relative L2 difference of phi and u between the
  two runs, and from the default run: phi and the three velocity
  components at `nsample` seeded grid points, the global L2 norms of phi
  and u, and the positions of every sphere (or of a seeded sample of them
  when `--pos-sample` is given).

tests/test_gpu_long_horizon.py runs the GPU step over the same workload and
compares with these checkpoints.  The workloads are the headline's sphere
density at 128^3 (12,500 spheres) and the headline itself (256^3, 100,000
spheres), both with radius 2 dx, gap 0.5 dx, contact angles U(60, 120) deg,
phi0 = U(-0.05, 0.05), dt = 0.01, FP64.

Usage: python tests/golden/make_long_horizon.py --n 128 --particles 12500
       [--steps 1000] [--every 50] [--threads 8] [--pos-sample 0]
       [--pos-steps 500,800,900,1000]

The committed fixtures were made with
  --n 128 --particles 12500 --steps 1000 --every 50 --pos-steps 500,800,900,1000
  --n 256 --particles 100000 --steps 1000 --every 100 --pos-steps 500,1000
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def workload(n, npart, steps):
    import fipi_amd
    p = fipi_amd.config_params(n, npart, 2.0, steps)
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, 1)
    pos = fipi_amd.init_rsa(npart, 2.0, 0.5, (p.lx, p.ly, p.lz), 2)
    assert pos.shape[0] == npart, pos.shape
    cth = np.cos(np.deg2rad(fipi_amd.init_uniform(npart, 60.0, 120.0, 3)))
    return p, phi0, pos, cth


def sample_points(G, nsample, seed=12345):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(G, size=nsample, replace=False)).astype(np.int64)


def pos_diff(a, b, L):
    d = a - b
    d -= L * np.round(d / L)
    return d


def main():
    import oracle
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--particles", type=int, default=12500)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--nsample", type=int, default=4096)
    ap.add_argument("--pos-sample", type=int, default=0,
                    help="store this many seeded spheres' positions (0: all)")
    ap.add_argument("--pos-every", type=int, default=0,
                    help="store positions at multiples of this step (0: every checkpoint)")
    ap.add_argument("--pos-steps", default="",
                    help="comma-separated checkpoint steps that store positions (overrides --pos-every)")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()

    n, npart, steps = args.n, args.particles, args.steps
    pos_steps_set = {int(x) for x in args.pos_steps.split(",") if x}
    p, phi0, pos, cth = workload(n, npart, steps)
    G = p.G
    L = np.array([p.lx, p.ly, p.lz])
    idx = sample_points(G, args.nsample)
    pidx = (np.arange(npart) if args.pos_sample <= 0 else
            np.sort(np.random.default_rng(777).choice(npart, size=args.pos_sample, replace=False)))
    oracle.set_fft_order(0)
    A = oracle.Sim(p, phi0, pos, None, cth, n_threads=args.threads)
    oracle.set_fft_order(1)
    B = oracle.Sim(p, phi0, pos, None, cth, n_threads=args.threads)
    ab_maxdp = np.zeros(steps)
    ab_maxid = np.zeros(steps, np.int32)
    ab_pos_rel = np.zeros(steps)
    ck_steps, ck_phi, ck_u, ck_phin, ck_un, ck_ab_phi, ck_ab_u = [], [], [], [], [], [], []
    pos_steps, pos_store = [], []
    pnorm = np.linalg.norm(pos)
    out = args.out or os.path.join(os.path.dirname(os.path.abspath(__file__)), f"long_{n}.npz")

    def save(t_done):
        # everything recorded through step t_done (rewritten at every
        # position checkpoint, so a long run leaves a usable fixture);
        # run_steps: the run length the workload was made for (the contact
        # ramp spans 0.2 of it)
        np.savez_compressed(
            out, n=n, particles=npart, steps=t_done, run_steps=steps, sample_idx=idx, pos_idx=pidx.astype(np.int32),
            ab_max_disp=ab_maxdp[:t_done], ab_max_sphere=ab_maxid[:t_done], ab_pos_rel=ab_pos_rel[:t_done],
            ck_steps=np.array(ck_steps, np.int32), ck_phi=np.array(ck_phi), ck_u=np.array(ck_u),
            ck_phi_norm=np.array(ck_phin), ck_u_norm=np.array(ck_un), ck_ab_phi=np.array(ck_ab_phi),
            ck_ab_u=np.array(ck_ab_u), pos_steps=np.array(pos_steps, np.int32), pos=np.array(pos_store),
            pos0_norm=pnorm)
        print("wrote", out, "through step", t_done, flush=True)

    t0 = time.time()
    for t in range(1, steps + 1):
        oracle.set_fft_order(0)
        A.step(1)
        oracle.set_fft_order(1)
        B.step(1)
        pa, pb = A.particles()["p"], B.particles()["p"]
        d = np.sqrt((pos_diff(pa, pb, L) ** 2).sum(1))
        ab_maxdp[t - 1] = d.max()
        ab_maxid[t - 1] = int(d.argmax())
        ab_pos_rel[t - 1] = np.sqrt((d ** 2).sum()) / np.linalg.norm(pa)
        if t % args.every == 0:
            phia, phib = A.phi().reshape(-1), B.phi().reshape(-1)
            ua, ub = A.u().reshape(3, -1), B.u().reshape(3, -1)
            ck_steps.append(t)
            ck_phi.append(phia[idx])
            ck_u.append(ua[:, idx])
            ck_phin.append(np.linalg.norm(phia))
            ck_un.append(np.linalg.norm(ua))
            ck_ab_phi.append(np.linalg.norm(phia - phib) / np.linalg.norm(phia))
            ck_ab_u.append(np.linalg.norm(ua - ub) / np.linalg.norm(ua))
            keep = (t in pos_steps_set) if pos_steps_set else (args.pos_every <= 0 or t % args.pos_every == 0)
            if keep:
                pos_steps.append(t)
                pos_store.append(pa[pidx])
                save(t)
            print(f"t={t} {time.time() - t0:.0f}s  A-B: phi {ck_ab_phi[-1]:.2e} u {ck_ab_u[-1]:.2e} "
                  f"pos rel {ab_pos_rel[t - 1]:.2e} max|dp| {ab_maxdp[t - 1]:.2e} (sphere {ab_maxid[t - 1]})",
                  flush=True)
    save(steps)


if __name__ == "__main__":
    main()
