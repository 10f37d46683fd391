"""Batched ensemble (fipi_ensemble_*): many boxes of mixed edge lengths
advanced together, one graph replay per step.

Every box must advance exactly as it does alone through the single-box
path (fipi_create / fipi_step), bit for bit: the slab kernels run the same
per-line arithmetic, the per-box reductions cover the same workgroups and
contact/spread sums keep the same order.  The single-box path itself is
checked against the oracle in test_gpu_step.py.  Each graph replay is
compared, with host synchronisation and reads between replays."""
import math

import numpy as np
import pytest

import fipi_amd
from fipi_amd import FP32_MIXED, FP64, config_params

pytestmark = pytest.mark.gpu

EDGES = [32, 48, 32, 64, 48, 32, 80, 96]
COUNTS = [40, 0, 25, 150, 60, 10, 0, 200]


def _case(prec, seed=1, edges=EDGES, counts=COUNTS):
    common = config_params(32, 0, 2.0, 50, precision=prec).replace(switch_particle=1)
    boxes = []
    for b, (n, k) in enumerate(zip(edges, counts)):
        mean = -0.3 + 0.6 * ((b * 37 + seed) % 11) / 10.0
        phi = fipi_amd.init_phi((n, n, n), mean, 0.05, seed * 100 + b)
        pos = ct = None
        if k:
            pos = fipi_amd.init_rsa(k, 2.0, 0.5, (n, n, n), seed * 1000 + b)
            assert pos.shape[0] == k
            ct = np.cos(np.deg2rad(fipi_amd.init_uniform(k, 60.0, 120.0, seed * 10000 + b)))
        boxes.append((n, k, phi, pos, ct))
    return common, boxes


def _single(common, n, k):
    return config_params(n, k, 2.0, 50, precision=common.precision)


def _ensemble(common, boxes):
    ens = fipi_amd.Ensemble(common, [b[0] for b in boxes], [b[1] for b in boxes])
    for i, (n, k, phi, pos, ct) in enumerate(boxes):
        ens.set_box(i, phi, pos, None, ct)
    return ens


@pytest.mark.parametrize("prec", [FP64, FP32_MIXED])
def test_boxes_match_single_box_path(gpu, prec):
    print(f"ensemble vs single boxes prec={prec}", flush=True)
    steps = 6
    common, boxes = _case(prec)
    ref = []
    for n, k, phi, pos, ct in boxes:
        sim = fipi_amd.Simulation(_single(common, n, k))
        sim.set_state(phi, pos, None, ct)
        traj = []
        for _ in range(steps):
            sim.step(1)
            traj.append((sim.phi().copy(), sim.particles()["p"].copy() if k else None))
        ref.append(traj)
        sim.close()
    ens = _ensemble(common, boxes)
    for s in range(steps):
        ens.step(1)  # s = 0: direct launches; s >= 1: replays of the two captured graphs
        for b, (n, k, _, _, _) in enumerate(boxes):
            phi_e = ens.phi(b)
            assert np.array_equal(phi_e, ref[b][s][0]), (s, b, np.abs(phi_e - ref[b][s][0]).max())
            if k:
                assert np.array_equal(ens.particles(b), ref[b][s][1]), (s, b)
    d = ens.diagnostics(3)
    assert d["t"] == steps
    ens.close()


def test_tuning(gpu):
    """Tile widths leave the results bit for bit; the tuner's y-pass FFT
    plan may change rounding, and its whole choice reproduces bit for bit
    when set on a fresh ensemble."""
    print("ensemble tuning", flush=True)
    common, boxes = _case(FP32_MIXED, seed=2, edges=[32, 64, 48], counts=[30, 100, 0])
    a = _ensemble(common, boxes)
    a.step(4)
    pa = [a.phi(b) for b in range(3)]
    a.close()
    w = _ensemble(common, boxes)
    for n in (32, 48, 64):
        w.set_tuning(n, {"y_tile_bytes": 64, "hub_tile_bytes": 64, "stokes_tile_bytes": 128,
                         "y_values_per_thread": 0})
    w.step(4)
    for b in range(3):
        assert np.array_equal(w.phi(b), pa[b]), b
    w.close()
    e = _ensemble(common, boxes)
    e.tune(2)
    chosen = {}
    for n in (32, 48, 64):
        t = e.tuning(n)
        assert t["y_tile_bytes"] in (64, 128) and t["hub_tile_bytes"] in (64, 128)
        assert t["stokes_tile_bytes"] in (64, 128)
        assert t["y_values_per_thread"] in ((8, 4) if n == 64 else (4,))
        chosen[n] = t
    assert e.tuning(96) == {"y_tile_bytes": 0, "hub_tile_bytes": 0, "stokes_tile_bytes": 0,
                            "y_values_per_thread": 0}
    e.step(4)
    pe = [e.phi(b) for b in range(3)]
    e.close()
    r = _ensemble(common, boxes)
    for n, t in chosen.items():
        r.set_tuning(n, t)
    r.step(4)
    for b in range(3):
        assert np.array_equal(r.phi(b), pe[b]), b
        assert np.linalg.norm(pe[b] - pa[b]) <= 1e-5 * np.linalg.norm(pa[b]), b
    r.close()


def test_mass_per_box(gpu):
    """Mean composition of every box is kept (FP64 k = 0 mode in the mixed
    path); boxes start at different means."""
    print("ensemble mass", flush=True)
    common, boxes = _case(FP32_MIXED, seed=3, edges=[32, 48, 32], counts=[0, 0, 0])
    common = common.replace(switch_ns=0)
    e = _ensemble(common, boxes)
    m0 = [float(np.mean(b[2])) for b in boxes]
    e.step(20)
    for b in range(3):
        assert abs(e.diagnostics(b)["mean_phi"] - m0[b]) < 1e-12, b
    e.close()


def test_ensemble_errors(gpu):
    common, _ = _case(FP64)
    with pytest.raises(fipi_amd.FipiError):
        fipi_amd.Ensemble(common, [40], [0])  # unsupported edge
    e = fipi_amd.Ensemble(common, [32, 32], [0, 0])
    e.set_box(0, np.zeros(32 ** 3))
    with pytest.raises(fipi_amd.FipiError):
        e.step(1)  # box 1 has no state
    e.close()
