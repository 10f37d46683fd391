"""Snapshot / restart (fipi_snapshot_*): a run continued from a snapshot
in a fresh context matches the uninterrupted run bit for bit, in both
precisions, with particles; and an FP64 run restored mid-way (inside the
SBDF2 history and the contact-force ramp) matches the oracle run straight
through."""
import numpy as np
import pytest

import fipi_amd
import oracle
from fipi_amd import FP32_MIXED, FP64
from _cases import make_case, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [FP64, FP32_MIXED])
def test_restart_is_exact(gpu, prec):
    print(f"snapshot prec={prec}", flush=True)
    p, phi0, pos = make_case(32, 120, steps=40, seed=9, precision=prec)
    a = fipi_amd.Simulation(p)
    a.set_state(phi0, pos)
    a.step(7)
    snap = a.snapshot()
    assert snap[:8] == b"FIPISNAP"
    a.step(9)
    ref_phi, ref_p, ref_t = a.phi(), a.particles()["p"], a.t
    a.close()
    b = fipi_amd.Simulation(p)
    b.restore(snap)
    assert b.t == 7
    b.step(9)
    assert b.t == ref_t
    assert np.array_equal(b.phi(), ref_phi)
    assert np.array_equal(b.particles()["p"], ref_p)
    b.close()


def test_restart_against_oracle(gpu):
    print("snapshot restart vs oracle", flush=True)
    p, phi0, pos = make_case(32, 200, steps=40, seed=19, gap=-0.5)
    cth = np.cos(np.deg2rad(fipi_amd.init_uniform(200, 60.0, 120.0, 20)))
    a = fipi_amd.Simulation(p)
    a.set_state(phi0, pos, None, cth)
    a.step(5)
    snap = a.snapshot()
    a.close()
    b = fipi_amd.Simulation(p)
    b.restore(snap)
    b.step(11)
    assert b.t == 16
    ref = oracle.Sim(p, phi0, pos, None, cth)
    ref.step(16)
    assert rel_l2(b.phi(), ref.phi()) < 1e-10
    assert rel_l2(b.velocity(), ref.u()) < 1e-10
    assert rel_l2(b.particles()["p"], ref.particles()["p"]) < 1e-10
    b.close()


def test_snapshot_rejects_other_grid(gpu):
    p, phi0, pos = make_case(32, 0, steps=10, seed=3)
    a = fipi_amd.Simulation(p)
    a.set_state(phi0, pos)
    snap = a.snapshot()
    a.close()
    q, _, _ = make_case(48, 0, steps=10, seed=3)
    b = fipi_amd.Simulation(q)
    with pytest.raises(fipi_amd.FipiError):
        b.restore(snap)
    with pytest.raises(fipi_amd.FipiError):
        b.restore(b"not a snapshot at all")
    b.close()


@pytest.mark.parametrize("prec", [FP64, FP32_MIXED])
def test_tune_keeps_state(gpu, prec):
    """fipi_tune times every choice and restores the state: the run
    continues exactly as a run that was given the same choice."""
    print(f"tune prec={prec}", flush=True)
    p, phi0, pos = make_case(32, 120, steps=40, seed=10, precision=prec)
    a = fipi_amd.Simulation(p)
    a.set_state(phi0, pos)
    a.step(3)
    chosen = a.tune(2)
    assert {chosen[k] for k in ("y_tile_bytes", "hub_tile_bytes", "stokes_tile_bytes")} <= {64, 128}
    assert chosen["y_values_per_thread"] == 4  # 32-point lines have one plan, 4.2.4
    assert a.tuning() == chosen
    assert a.t == 3
    a.step(6)
    b = fipi_amd.Simulation(p)
    b.set_state(phi0, pos)
    b.set_tuning(chosen)
    b.step(3)
    b.step(6)
    assert np.array_equal(a.phi(), b.phi())
    assert np.array_equal(a.particles()["p"], b.particles()["p"])
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [FP64, FP32_MIXED])
def test_one_call_matches_single_steps(gpu, prec):
    """Within one fipi_step call each step's particle stage runs on a second
    stream beside the next step's field passes; step-by-step calls run it in
    order.  Both give the same state bit for bit."""
    print(f"overlap prec={prec}", flush=True)
    p, phi0, pos = make_case(32, 150, steps=40, seed=12, precision=prec)
    a = fipi_amd.Simulation(p)
    a.set_state(phi0, pos)
    a.step(12)
    b = fipi_amd.Simulation(p)
    b.set_state(phi0, pos)
    for _ in range(12):
        b.step(1)
    assert np.array_equal(a.phi(), b.phi())
    assert np.array_equal(a.grad_phi(), b.grad_phi())
    assert np.array_equal(a.particles()["p"], b.particles()["p"])
    a.close()
    b.close()
