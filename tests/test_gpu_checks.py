"""GPU step checks against the oracle and against independent formulas:

* diagnostics: free energy sum lambda(|grad phi|^2/2 + (phi^2-1)^2/(4 eps^2))
  dx^3 (main.cpp:232, field.cpp:506-516), interface area sum |grad phi| dx^3
  and mean phi against a NumPy evaluation on the fetched phi with its own
  spectral gradient; the packing energy tau (cluster.cpp:256-259) against
  the oracle's;
* buoyancy with rho1 != rho2 (head.hpp:44,53; main.cpp:172), which only
  matters when the density contrast makes the force non-uniform;
* every y-pass FFT plan of 64-, 128- and 256-point lines inside the full
  step (CH, Stokes, spheres) on grids long in y, FP64 against the oracle
  and FP32-mixed plan against plan;
* FP32-mixed against the FP64 oracle at the headline's sphere density.
"""
import os

import numpy as np
import pytest

import fipi_amd
import oracle
from _cases import make_case, rel_l2
from fipi_amd import config_params

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _spectral_grad(c, p):
    nz, ny, nx = c.shape
    ch = np.fft.fftn(c)
    out = []
    for axis, (N, L) in enumerate(((nx, p.lx), (ny, p.ly), (nz, p.lz))):
        idx = np.arange(N)
        k = np.where(idx <= N // 2, idx, idx - N) * 2.0 * p.pi / L
        k[N // 2] = 0.0
        shape = [1, 1, 1]
        shape[2 - axis] = N
        out.append(np.real(np.fft.ifftn(1j * k.reshape(shape) * ch)))
    return np.array(out)


def test_diagnostics_against_numpy_and_oracle_tau(gpu):
    print("diagnostics vs numpy, tau vs oracle", flush=True)
    # overlapping spheres on a small-amplitude field: contacts and f_pi != 0
    p, phi0, pos = make_case(64, 2500, steps=50, seed=31, gap=-1.0)
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0, pos)
    sim.step(6)
    phi, g = sim.phi(), sim.grad_phi()
    d = sim.diagnostics()
    sim.close()
    gn = _spectral_grad(phi, p)
    assert rel_l2(g, gn) < 1e-12
    dx3 = p.dx ** 3
    gg = (gn ** 2).sum(0)
    fe = float((p.lam * (0.5 * gg + 0.25 / p.epsilon ** 2 * (phi ** 2 - 1.0) ** 2)).sum() * dx3)
    area = float(np.sqrt(gg).sum() * dx3)
    assert abs(d["free_energy"] / fe - 1) < 1e-12, (d["free_energy"], fe)
    assert abs(d["interface_area"] / area - 1) < 1e-12, (d["interface_area"], area)
    assert abs(d["mean_phi"] - phi.mean()) < 1e-15
    ref = oracle.Sim(p, phi0, pos)
    ref.step(6)
    tau_ref = ref.tau()
    assert tau_ref != 0.0
    assert abs(d["tau"] / tau_ref - 1) < 1e-10, (d["tau"], tau_ref)


def test_buoyancy_contrast(gpu):
    """rho2 != rho1: ((1+c)/2 rho1 + (1-c)/2 rho2) g varies with c, so the
    Stokes flow has a buoyancy-driven part."""
    print("buoyancy rho1 != rho2", flush=True)
    p, phi0, pos = make_case(32, 150, steps=40, seed=32, amp=0.3)
    p = p.replace(rho1=1.0, rho2=0.4, gravity=-2.0)
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0, pos)
    sim.step(20)
    ref = oracle.Sim(p, phi0, pos)
    ref.step(20)
    u_ref = ref.u()
    assert rel_l2(sim.phi(), ref.phi()) < TOL
    assert rel_l2(sim.velocity(), u_ref) < TOL
    assert rel_l2(sim.particles()["p"], ref.particles()["p"]) < TOL
    # the contrast matters: without it the flow differs at O(1)
    flat = oracle.Sim(p.replace(rho2=1.0), phi0, pos)
    flat.step(20)
    assert rel_l2(flat.u(), u_ref) > 1e-2
    sim.close()


@pytest.mark.parametrize("ny", [64, 128, 256])
def test_y_plans_in_the_step(gpu, ny):
    """Both y-pass plans of a line length inside the full step (the
    gradient, one- and three-field y passes), on a 32 x ny x 32 grid."""
    print(f"y plans in the step, ny = {ny}", flush=True)
    n = 32
    p = config_params(n, 0, 2.0, 40)
    p = p.replace(ny=ny, ly=float(ny), n_particles=400, switch_particle=1)
    phi0 = fipi_amd.init_phi((n, ny, n), 0.0, 0.05, 41)
    pos = fipi_amd.init_rsa(400, 2.0, 0.5, (p.lx, p.ly, p.lz), 42)
    assert pos.shape[0] == 400
    ref = oracle.Sim(p, phi0, pos)
    ref.step(8)
    r_phi, r_u, r_p = ref.phi(), ref.u(), ref.particles()["p"]
    for prec in (fipi_amd.FP64, fipi_amd.FP32_MIXED):
        q = p.replace(precision=prec)
        outs = {}
        probe = fipi_amd.Simulation(q)
        probe.set_state(phi0, pos)
        plans = {probe.tuning()["y_values_per_thread"]}
        # the other plan of this length
        for e in (4, 8, 16):
            t = probe.tuning()
            t["y_values_per_thread"] = e
            try:
                probe.set_tuning(t)
                plans.add(e)
            except fipi_amd.FipiError:
                pass
        probe.close()
        assert len(plans) == 2, plans
        for e in sorted(plans):
            sim = fipi_amd.Simulation(q)
            sim.set_state(phi0, pos)
            t = sim.tuning()
            t["y_values_per_thread"] = e
            sim.set_tuning(t)
            sim.step(8)
            outs[e] = (sim.phi(), sim.velocity(), sim.particles()["p"])
            sim.close()
        for e, (phi, u, pp) in outs.items():
            if prec == fipi_amd.FP64:
                assert rel_l2(phi, r_phi) < TOL, (e, rel_l2(phi, r_phi))
                assert rel_l2(u, r_u) < TOL, (e, rel_l2(u, r_u))
                assert rel_l2(pp, r_p) < TOL, (e, rel_l2(pp, r_p))
            else:
                assert rel_l2(phi, r_phi) < 1e-5, (e, rel_l2(phi, r_phi))
        if prec == fipi_amd.FP32_MIXED:
            a, b = (outs[e] for e in sorted(outs))
            assert rel_l2(a[0], b[0]) < 1e-5


def test_mixed_vs_oracle_headline_density_128(gpu):
    """FP32-mixed against the FP64 oracle (not the FP64 GPU path): 128^3,
    12,500 spheres of radius 2 dx, contact angles U(60, 120) deg, 200 steps;
    phi, interface area and sphere positions."""
    print("mixed vs oracle 128^3 + 12.5k, 200 steps", flush=True)
    n, npart, steps = 128, 12500, 200
    p = fipi_amd.config_params(n, npart, 2.0, steps)
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, 1)
    pos = fipi_amd.init_rsa(npart, 2.0, 0.5, (p.lx, p.ly, p.lz), 2)
    cth = np.cos(np.deg2rad(fipi_amd.init_uniform(npart, 60.0, 120.0, 3)))
    sim = fipi_amd.Simulation(p.replace(precision=fipi_amd.FP32_MIXED))
    sim.set_state(phi0, pos, None, cth)
    sim.step(steps)
    m_phi, m_p = sim.phi(), sim.particles()["p"]
    sim.close()
    threads = min(len(os.sched_getaffinity(0)), 16)
    ref = oracle.Sim(p, phi0, pos, None, cth, n_threads=threads)
    ref.step(steps)
    r_phi, r_p = ref.phi(), ref.particles()["p"]
    e_phi = rel_l2(m_phi, r_phi)
    area = lambda c: float(np.sqrt((_spectral_grad(c, p) ** 2).sum(0)).sum())
    e_area = abs(area(m_phi) / area(r_phi) - 1)
    L = np.array([p.lx, p.ly, p.lz])
    dp = m_p - r_p
    dp -= L * np.round(dp / L)
    print(f"mixed vs oracle: phi {e_phi:.2e} area {e_area:.2e} max |dp| {np.abs(dp).max():.2e}", flush=True)
    assert e_phi < 1e-5
    assert e_area < 1e-5
    assert np.abs(dp).max() < 1e-3


@pytest.mark.parametrize("prec", [fipi_amd.FP64, fipi_amd.FP32_MIXED])
def test_tile_widths_leave_results_unchanged(gpu, prec):
    """Every combination of y-pass, CH-hub and Stokes-hub column tiles (64
    or 128 bytes; 64-byte tiles run in XCD-paired order) gives bitwise the
    same step, and FP64 matches the oracle; a grid with 48 rows exercises
    an odd tile count per row."""
    print(f"tile widths prec={prec}", flush=True)
    for n in (64, 48):
        p, phi0, pos = make_case(n, 300 if n == 64 else 150, steps=40, seed=71, precision=prec)
        outs = []
        for y in (64, 128):
            for h in (64, 128):
                for s in (64, 128):
                    sim = fipi_amd.Simulation(p)
                    sim.set_state(phi0, pos)
                    t = sim.tuning()
                    t.update(y_tile_bytes=y, hub_tile_bytes=h, stokes_tile_bytes=s)
                    sim.set_tuning(t)
                    sim.step(6)
                    outs.append((sim.phi(), sim.velocity(), sim.particles()["p"]))
                    sim.close()
        for o in outs[1:]:
            for a, b in zip(o, outs[0]):
                assert np.array_equal(a, b)
        if prec == fipi_amd.FP64:
            ref = oracle.Sim(p, phi0, pos)
            ref.step(6)
            assert rel_l2(outs[0][0], ref.phi()) < TOL
            assert rel_l2(outs[0][1], ref.u()) < TOL
            assert rel_l2(outs[0][2], ref.particles()["p"]) < TOL
