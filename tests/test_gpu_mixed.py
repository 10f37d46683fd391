"""FP32-mixed path: FFTs, k-space update, interpolation and spreading in
FP32; the k = 0 (mass) mode, contact overlaps and particle positions in
FP64.  Checked against the FP64 GPU path, which test_gpu_step.py pins to
the oracle at 1e-10."""
import numpy as np
import pytest

import fipi_amd
from _cases import make_case, rel_l2

pytestmark = pytest.mark.gpu


def _run(p, phi0, pos, cth, steps):
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0, pos, None, cth)
    sim.step(steps)
    out = sim.phi(), sim.particles()["p"], sim.diagnostics()
    sim.close()
    return out


def test_mixed_vs_fp64_64(gpu):
    print("mixed vs fp64 64^3", flush=True)
    steps = 200
    p64, phi0, pos = make_case(64, 1500, steps=steps, seed=21, mean=0.1)
    cth = np.cos(np.deg2rad(fipi_amd.init_uniform(1500, 60.0, 120.0, 5)))
    pmx = p64.replace(precision=fipi_amd.FP32_MIXED)
    a_phi, a_p, a_d = _run(p64, phi0, pos, cth, steps)
    b_phi, b_p, b_d = _run(pmx, phi0, pos, cth, steps)
    e = rel_l2(b_phi, a_phi)
    print(f"rel L2 phi {e:.2e}, max |dp| {np.abs(b_p - a_p).max():.2e}, "
          f"area {abs(b_d['interface_area'] / a_d['interface_area'] - 1):.2e}")
    assert e < 1e-5
    assert np.abs(b_p - a_p).max() < 1e-4
    assert abs(b_d["interface_area"] / a_d["interface_area"] - 1) < 1e-5
    # the mass (k = 0 mode) is carried in FP64
    assert abs(b_d["mean_phi"] - a_d["mean_phi"]) < 1e-12


def test_mixed_mass_conservation(gpu):
    """Without flow conv = 0 and the SBDF2 update keeps the k = 0 mode fixed
    (main.cpp:117-127 with k^2 = 0); the FP64 mass mode holds it to 1e-12
    where FP32 storage of the mean alone would drift by ~1e-8.  (With flow
    the reference's own mass moves through conv_hat(0), since its
    projection is not solenoidal on the Nyquist planes; that case is
    compared with the FP64 path above.)"""
    p, phi0, _ = make_case(64, 0, steps=100, seed=22, mean=0.3, amp=0.1,
                           precision=fipi_amd.FP32_MIXED, ns=0)
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0)
    m0 = phi0.mean()
    sim.step(100)
    d = sim.diagnostics()
    assert abs(d["mean_phi"] - m0) < 1e-12
    sim.close()


def test_mixed_vs_fp64_headline_256(gpu):
    """The headline workload (256^3, 100,000 spheres of radius 2 dx, contact
    angles U(60, 120) deg), 300 steps in each precision from the same seeded
    state: the FP32-mixed tolerance (1e-5 relative L2 in phi) at full size."""
    print("mixed vs fp64 256^3 + 100k, 300 steps", flush=True)
    n, npart, steps = 256, 100000, 300
    p64 = fipi_amd.config_params(n, npart, 2.0, steps)
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, 1)
    pos = fipi_amd.init_rsa(npart, 2.0, 0.5, (p64.lx, p64.ly, p64.lz), 2)
    assert pos.shape[0] == npart
    cth = np.cos(np.deg2rad(fipi_amd.init_uniform(npart, 60.0, 120.0, 3)))
    a_phi, a_p, a_d = _run(p64, phi0, pos, cth, steps)
    b_phi, b_p, b_d = _run(p64.replace(precision=fipi_amd.FP32_MIXED), phi0, pos, cth, steps)
    e = rel_l2(b_phi, a_phi)
    dp = np.abs(b_p - a_p)
    print(f"rel L2 phi {e:.2e}, max |dp| {dp.max():.2e}", flush=True)
    assert e < 1e-5
    # a sphere on a near-zero grad c turns its capillary force with the
    # round-off, so the bound is on the largest displacement, in dx
    assert dp.max() < 1e-3
    assert abs(b_d["interface_area"] / a_d["interface_area"] - 1) < 1e-5
    assert abs(b_d["mean_phi"] - a_d["mean_phi"]) < 1e-12
    # every sphere stays in the box
    L = np.array([p64.lx, p64.ly, p64.lz])
    assert np.all((a_p >= 0) & (a_p < L)) and np.all((b_p >= 0) & (b_p < L))
