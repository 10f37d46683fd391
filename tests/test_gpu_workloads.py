"""The batched ensemble and the calibration adjoint at the workloads the
benchmark quotes them on.

* Ensemble: 512 boxes with edges from {32, 48, 64, 80, 96}, mean
  composition U(-0.3, 0.3), U(0, 2000) spheres per box capped where random
  sequential addition jams, FP32-mixed, the first step plus 20 replays of
  the captured step graph.  16 sampled boxes equal the same boxes stepped
  alone through the single-box path bit for bit, and their FP64 mass mode
  equals the FP64 path's mean at 1e-12; the three first 32^3 boxes agree
  with the FP64 oracle to the FP32-mixed bar.
* Adjoint: 128^3 with 20,000 spheres of radius 2 dx held fixed, 200 steps
  with checkpoints every 20, target from M = 1.3, kappa = 0.8 and contact
  angles U(70, 110) deg; gradients in M, kappa and the three most sensitive
  contact angles against FP64 central differences (h = 1e-4) at 1e-6.
"""
import os
import sys

import numpy as np
import pytest

import fipi_amd
import oracle
from _cases import rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_ensemble_512_boxes(gpu):
    print("ensemble 512 boxes, 21 steps", flush=True)
    sys.path.insert(0, ROOT)
    from bench import ensemble_workload
    boxes, _, _ = ensemble_workload(512)
    steps = 21
    common = fipi_amd.config_params(32, 0, 2.0, 500, precision=fipi_amd.FP32_MIXED).replace(switch_particle=1)
    ens = fipi_amd.Ensemble(common, [b[0] for b in boxes], [b[1] for b in boxes])
    for i, (n, k, phi, pos) in enumerate(boxes):
        ens.set_box(i, phi, pos if k else None)
    ens.step(steps)
    # the mass moves only through conv_hat(0) (the reference's projection is
    # not solenoidal on the Nyquist planes), by a few 1e-9 in 21 steps
    means = [ens.diagnostics(b)["mean_phi"] for b in range(len(boxes))]
    drift = max(abs(means[b] - boxes[b][2].mean()) for b in range(len(boxes)))
    print(f"largest mass drift {drift:.2e}", flush=True)
    assert drift < 1e-7
    rng = np.random.default_rng(5)
    sample = sorted(rng.choice(len(boxes), size=16, replace=False).tolist())
    ens_phi = {b: ens.phi(b) for b in sample}
    ens_pos = {b: ens.particles(b) for b in sample if boxes[b][1]}
    small = [b for b in range(len(boxes)) if boxes[b][0] == 32][:3]
    small_phi = {b: ens.phi(b) for b in small}
    ens.close()
    for b in sample:
        n, k, phi, pos = boxes[b]
        sim = fipi_amd.Simulation(fipi_amd.config_params(n, k, 2.0, 500, precision=fipi_amd.FP32_MIXED))
        sim.set_state(phi, pos if k else None)
        sim.step(steps)
        assert np.array_equal(sim.phi(), ens_phi[b]), b
        if k:
            assert np.array_equal(sim.particles()["p"], ens_pos[b]), b
        sim.close()
        # the FP64 mass mode of the mixed path against the FP64 path
        s64 = fipi_amd.Simulation(fipi_amd.config_params(n, k, 2.0, 500))
        s64.set_state(phi, pos if k else None)
        s64.step(steps)
        m64 = s64.diagnostics()["mean_phi"]
        s64.close()
        assert abs(means[b] - m64) < 1e-12, (b, means[b], m64)
    for b in small:
        n, k, phi, pos = boxes[b]
        p = fipi_amd.config_params(n, k, 2.0, 500)
        ref = oracle.Sim(p, phi, pos if k else None)
        ref.step(steps)
        e = rel_l2(small_phi[b], ref.phi())
        print(f"box {b} ({n}^3, {k} spheres) vs oracle: {e:.2e}", flush=True)
        assert e < 1e-5


def test_adjoint_128_20k(gpu):
    print("adjoint 128^3 + 20k, 200 steps", flush=True)
    n, npart, steps, every, seed = 128, 20000, 200, 20, 21
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, seed)
    pos = fipi_amd.init_rsa(npart, 2.0, 0.0, (n, n, n), seed + 1)
    assert pos.shape[0] == npart
    th_true = np.deg2rad(fipi_amd.init_uniform(npart, 70.0, 110.0, seed + 2))
    th0 = np.full(npart, np.pi / 2)
    p = fipi_amd.config_params(n, npart, 2.0, steps)
    adj = fipi_amd.Adjoint(p, phi0, pos, 0.5, steps, every)
    _, target = adj.forward(1.3, 0.8, th_true, want_loss=False)
    adj.set_target(target)
    loss, gm, gk, gth = adj.gradient(1.0, 1.0, th0)
    assert loss > 0
    h = 1e-4

    def f(mm, kk, tt):
        return adj.forward(mm, kk, tt)[0]
    fm = (f(1.0 + h, 1.0, th0) - f(1.0 - h, 1.0, th0)) / (2 * h)
    fk = (f(1.0, 1.0 + h, th0) - f(1.0, 1.0 - h, th0)) / (2 * h)
    print(f"dL/dM {gm:.10e} fd {fm:.10e}; dL/dkappa {gk:.10e} fd {fk:.10e}", flush=True)
    assert abs(gm - fm) <= 1e-6 * abs(fm)
    assert abs(gk - fk) <= 1e-6 * abs(fk)
    gs = np.abs(gth).max()
    assert gs > 0
    for i in np.argsort(-np.abs(gth))[:3]:
        tp, tm = th0.copy(), th0.copy()
        tp[i] += h
        tm[i] -= h
        fd = (f(1.0, 1.0, tp) - f(1.0, 1.0, tm)) / (2 * h)
        print(f"dL/dtheta[{i}] {gth[i]:.10e} fd {fd:.10e}", flush=True)
        assert abs(gth[i] - fd) <= 1e-6 * gs
    adj.close()
