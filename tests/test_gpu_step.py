"""FP64 GPU time step against the CPU oracle (restatement of
src/main.cpp:75-222) on identical seeded inputs.

Bar: relative L2 error of phi, u and particle positions <= 1e-10 (the
FP64 tolerance of the path); discrete sets (cells, contact pairs) bit-exact
in test_gpu_particles.py."""
import numpy as np
import pytest

import fipi_amd
import oracle
from _cases import make_case, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _run_both(p, phi0, pos, steps, chunks=1):
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0, pos)
    ref = oracle.Sim(p, phi0, pos)
    per = steps // chunks
    out = []
    for _ in range(chunks):
        sim.step(per)
        ref.step(per)
        out.append((sim.t, sim.phi(), ref.phi(), sim.velocity(), ref.u(),
                    sim.particles() if sim.n else None, ref.particles() if sim.n else None))
    sim.close()
    return out


@pytest.mark.parametrize("ns", [0, 1])
def test_ch_only_32(gpu, ns):
    print(f"ch-only ns={ns}", flush=True)
    p, phi0, _ = make_case(32, 0, steps=20, ns=ns, seed=3)
    for t, g_phi, r_phi, g_u, r_u, _, _ in _run_both(p, phi0, None, 20, chunks=4):
        e = rel_l2(g_phi, r_phi)
        assert e < TOL, (t, e)
        if ns:
            assert rel_l2(g_u, r_u) < TOL, (t, rel_l2(g_u, r_u))


def test_step_with_particles_32(gpu):
    print("particles 32^3", flush=True)
    p, phi0, pos = make_case(32, 150, steps=40, seed=5)
    for t, g_phi, r_phi, g_u, r_u, gp, rp in _run_both(p, phi0, pos, 40, chunks=4):
        assert rel_l2(g_phi, r_phi) < TOL, (t, rel_l2(g_phi, r_phi))
        assert rel_l2(g_u, r_u) < TOL, (t, rel_l2(g_u, r_u))
        assert rel_l2(gp["p"], rp["p"]) < TOL, (t, rel_l2(gp["p"], rp["p"]))
        assert rel_l2(gp["f_pp"], rp["f_pp"]) < 1e-8, (t, rel_l2(gp["f_pp"], rp["f_pp"]))
        assert rel_l2(gp["f_pi"], rp["f_pi"]) < 1e-8, (t, rel_l2(gp["f_pi"], rp["f_pi"]))


@pytest.mark.parametrize("prec", [fipi_amd.FP64, fipi_amd.FP32_MIXED])
def test_particles_without_stokes_32(gpu, prec):
    """switch_ns = 0 with particles: no force pass, so within one call the
    particle stage of each step reads c and grad c from the set the next
    step does not write (the spare set).  FP64 against the oracle; in both
    precisions one call against single steps, bitwise."""
    print(f"particles ns=0 prec={prec}", flush=True)
    p, phi0, pos = make_case(32, 150, steps=40, seed=6, ns=0, precision=prec)
    if prec == fipi_amd.FP64:
        for t, g_phi, r_phi, _, _, gp, rp in _run_both(p, phi0, pos, 12, chunks=2):
            assert rel_l2(g_phi, r_phi) < TOL, (t, rel_l2(g_phi, r_phi))
            assert rel_l2(gp["p"], rp["p"]) < TOL, (t, rel_l2(gp["p"], rp["p"]))
            assert rel_l2(gp["f_pi"], rp["f_pi"]) < 1e-8, (t, rel_l2(gp["f_pi"], rp["f_pi"]))
    a = fipi_amd.Simulation(p)
    a.set_state(phi0, pos)
    a.step(12)
    b = fipi_amd.Simulation(p)
    b.set_state(phi0, pos)
    for _ in range(12):
        b.step(1)
    assert np.array_equal(a.phi(), b.phi())
    assert np.array_equal(a.particles()["p"], b.particles()["p"])
    a.close()
    b.close()


@pytest.mark.parametrize("n,npart", [(48, 120), (80, 300), (96, 400)])
def test_step_odd_radix(gpu, n, npart):
    """Edges with radix-3/5/6 transforms (the ensemble's 48, 80, 96)."""
    print(f"odd-radix {n}^3 + {npart}", flush=True)
    p, phi0, pos = make_case(n, npart, steps=30, seed=n)
    for t, g_phi, r_phi, g_u, r_u, gp, rp in _run_both(p, phi0, pos, 6, chunks=2):
        assert rel_l2(g_phi, r_phi) < TOL, (t, rel_l2(g_phi, r_phi))
        assert rel_l2(g_u, r_u) < TOL, (t, rel_l2(g_u, r_u))
        assert rel_l2(gp["p"], rp["p"]) < TOL, (t, rel_l2(gp["p"], rp["p"]))


def test_step_config1_64(gpu):
    """64^3 benchmark units with 2.5 dx spheres (config 1: 2,000 requested,
    capped at the ~1,060 random sequential addition places), 10 steps."""
    print("config1 64^3", flush=True)
    p, phi0, pos = make_case(64, 2000, r_p=2.5, steps=200, seed=7, cap=True)
    (t, g_phi, r_phi, g_u, r_u, gp, rp), = _run_both(p, phi0, pos, 10)
    assert rel_l2(g_phi, r_phi) < TOL
    assert rel_l2(g_u, r_u) < TOL
    assert rel_l2(gp["p"], rp["p"]) < TOL


def test_1000_steps_64(gpu):
    """The FP64 bar (1e-10 relative L2 after 1000 steps) at 64^3
    with config 1's spheres of radius 2.5 (2,000 requested, capped at what
    random sequential addition places), checked every 250 steps; the oracle
    takes about half a minute for this on the host."""
    print("1000 steps 64^3", flush=True)
    p, phi0, pos = make_case(64, 2000, r_p=2.5, steps=1000, seed=7, cap=True)
    for t, g_phi, r_phi, g_u, r_u, gp, rp in _run_both(p, phi0, pos, 1000, chunks=4):
        print(t, rel_l2(g_phi, r_phi), rel_l2(g_u, r_u), rel_l2(gp["p"], rp["p"]), flush=True)
        assert rel_l2(g_phi, r_phi) < TOL, (t, rel_l2(g_phi, r_phi))
        assert rel_l2(g_u, r_u) < TOL, (t, rel_l2(g_u, r_u))
        assert rel_l2(gp["p"], rp["p"]) < TOL, (t, rel_l2(gp["p"], rp["p"]))


def test_reference_defaults_64(gpu):
    """The reference's own constants (head.hpp: L = 2 pi, dt = 1e-3,
    r_eq = 5 r_p), 400 particles, 5 steps."""
    print("reference defaults", flush=True)
    p = fipi_amd.reference_params(n_particles=400)
    phi0 = fipi_amd.init_phi((64, 64, 64), 0.0, 0.1, 11)
    pos = fipi_amd.init_rsa(400, p.r_p, 0.0, (p.lx, p.ly, p.lz), 12)
    assert pos.shape[0] == 400
    (t, g_phi, r_phi, g_u, r_u, gp, rp), = _run_both(p, phi0, pos, 5)
    assert rel_l2(g_phi, r_phi) < TOL
    assert rel_l2(g_u, r_u) < TOL
    assert rel_l2(gp["p"], rp["p"]) < TOL


def test_determinism(gpu):
    p, phi0, pos = make_case(32, 150, steps=20, seed=9)
    outs = []
    for _ in range(2):
        sim = fipi_amd.Simulation(p)
        sim.set_state(phi0, pos)
        sim.step(10)
        outs.append((sim.phi(), sim.velocity(), sim.particles()["p"]))
        sim.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


def test_headline_size_256(gpu):
    """The headline workload itself: 256^3 with 100,000 spheres of radius
    2 dx (gap 0.5 dx) and contact angles U(60, 120) deg, FP64, 3 steps
    against the oracle (the first-order start and two SBDF2 steps; both
    Heun stages and the spread each step)."""
    import os
    print("headline 256^3 + 100k, 3 steps", flush=True)
    n, npart = 256, 100000
    p = fipi_amd.config_params(n, npart, 2.0, 1000)
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, 1)
    pos = fipi_amd.init_rsa(npart, 2.0, 0.5, (p.lx, p.ly, p.lz), 2)
    assert pos.shape[0] == npart
    cth = np.cos(np.deg2rad(fipi_amd.init_uniform(npart, 60.0, 120.0, 3)))
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0, pos, None, cth)
    sim.step(3)
    g_phi, g_u, g_p = sim.phi(), sim.velocity(), sim.particles()
    sim.close()
    threads = min(len(os.sched_getaffinity(0)), 16)
    ref = oracle.Sim(p, phi0, pos, None, cth, n_threads=threads)
    ref.step(3)
    r_p = ref.particles()
    assert rel_l2(g_phi, ref.phi()) < TOL, rel_l2(g_phi, ref.phi())
    assert rel_l2(g_u, ref.u()) < TOL, rel_l2(g_u, ref.u())
    assert rel_l2(g_p["p"], r_p["p"]) < TOL, rel_l2(g_p["p"], r_p["p"])
    assert rel_l2(g_p["f_pi"], r_p["f_pi"]) < 1e-8, rel_l2(g_p["f_pi"], r_p["f_pi"])



def test_headline_density_128(gpu):
    """The headline's sphere density and contact angles at 128^3 (12,500
    spheres of radius 2 dx, U(60, 120) deg), 200 FP64 steps against the
    oracle (about half a minute on 16 host threads)."""
    import os
    print("128^3 + 12.5k, 200 steps", flush=True)
    n, npart, steps = 128, 12500, 200
    p = fipi_amd.config_params(n, npart, 2.0, steps)
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, 1)
    pos = fipi_amd.init_rsa(npart, 2.0, 0.5, (p.lx, p.ly, p.lz), 2)
    assert pos.shape[0] == npart
    cth = np.cos(np.deg2rad(fipi_amd.init_uniform(npart, 60.0, 120.0, 3)))
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0, pos, None, cth)
    threads = min(len(os.sched_getaffinity(0)), 16)
    ref = oracle.Sim(p, phi0, pos, None, cth, n_threads=threads)
    for _ in range(2):
        sim.step(steps // 2)
        ref.step(steps // 2)
        g_p, r_p = sim.particles(), ref.particles()
        e = (rel_l2(sim.phi(), ref.phi()), rel_l2(sim.velocity(), ref.u()), rel_l2(g_p["p"], r_p["p"]),
             rel_l2(g_p["f_pi"], r_p["f_pi"]))
        print(sim.t, e, flush=True)
        assert e[0] < TOL and e[1] < TOL and e[2] < TOL, (sim.t, e)
        assert e[3] < 1e-8, (sim.t, e)
    sim.close()


def test_particles_leaving_their_grid_cell_32(gpu):
    """Stiff contacts between overlapping spheres (k_n = 20, dt = 0.05):
    spheres cross grid cells within Heun stage 1, so stage 2 cannot reuse
    the corner values stage 1 gathered and takes them from the fields.  The
    oracle run shows that such crossings happen; FP64 against the oracle
    every step, tau included."""
    print("grid-cell crossings between the stages", flush=True)
    p, phi0, _ = make_case(32, 150, steps=20, seed=7)
    p = p.replace(k_n=20.0, dt=0.05)
    pos = fipi_amd.init_rsa(150, 2.0, -1.0, (p.lx, p.ly, p.lz), 7)
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0, pos)
    ref = oracle.Sim(p, phi0, pos)
    L = np.array([p.lx, p.ly, p.lz])
    crossings = 0
    for t in range(10):
        sim.step(1)
        ref.step(1)
        rp, gp = ref.particles(), sim.particles()
        p1 = np.fmod(rp["p_old"] + p.dt * rp["u_old"] + L, L)
        crossings += int(np.any(np.floor(rp["p_old"]) != np.floor(p1), axis=1).sum())
        assert rel_l2(sim.phi(), ref.phi()) < TOL, t
        assert rel_l2(gp["p"], rp["p"]) < TOL, (t, rel_l2(gp["p"], rp["p"]))
        assert rel_l2(gp["f_pi"], rp["f_pi"]) < 1e-8, t
        assert rel_l2(gp["f_pp"], rp["f_pp"]) < 1e-8, t
        tau_g, tau_r = sim.diagnostics()["tau"], ref.tau()
        assert abs(tau_g - tau_r) <= 1e-10 * max(1.0, abs(tau_r)), (t, tau_g, tau_r)
    sim.close()
    assert crossings > 0


@pytest.mark.parametrize("n,npart", [(32, 2000), (64, 12000)])
def test_crowded_cells_32_64(gpu, n, npart):
    """Small overlapping spheres (r = 0.5, contact distance 1) crowded into
    6-dx contact cells: 16 per cell on average at 32^3, so cells exceed the
    table's 16 slots and tiles take the exact path from global memory; at
    64^3 fast and exact tiles mix.  FP64 against the oracle."""
    print(f"crowded cells {n}^3 + {npart}", flush=True)
    p, phi0, _ = make_case(n, npart, r_p=0.5, steps=20, seed=41)
    pos = fipi_amd.init_rsa(npart, 0.5, -0.3, (p.lx, p.ly, p.lz), 41)
    assert pos.shape[0] == npart
    for t, g_phi, r_phi, g_u, r_u, gp, rp in _run_both(p, phi0, pos, 6, chunks=3):
        assert rel_l2(g_phi, r_phi) < TOL, (t, rel_l2(g_phi, r_phi))
        assert rel_l2(g_u, r_u) < TOL, (t, rel_l2(g_u, r_u))
        assert rel_l2(gp["p"], rp["p"]) < TOL, (t, rel_l2(gp["p"], rp["p"]))
        assert rel_l2(gp["f_pp"], rp["f_pp"]) < 1e-8, (t, rel_l2(gp["f_pp"], rp["f_pp"]))


def test_heavily_crowded_cells_32(gpu):
    """8,000 spheres of radius 0.5 in 6-dx contact cells: 64 a cell on
    average, four times the table's 16 slots, so every tile takes the exact
    path and most records sit in the cells' overflow chains.  FP64 against
    the oracle every two steps."""
    print("heavily crowded cells 32^3 + 8000, 6-dx cells", flush=True)
    n, npart = 32, 8000
    p, phi0, pos = make_case(n, npart, r_p=0.5, steps=20, seed=47, gap=-0.3)
    for t, g_phi, r_phi, g_u, r_u, gp, rp in _run_both(p, phi0, pos, 4, chunks=2):
        print(f"t={t}", flush=True)
        assert rel_l2(g_phi, r_phi) < TOL, (t, rel_l2(g_phi, r_phi))
        assert rel_l2(g_u, r_u) < TOL, (t, rel_l2(g_u, r_u))
        assert rel_l2(gp["p"], rp["p"]) < TOL, (t, rel_l2(gp["p"], rp["p"]))
        assert rel_l2(gp["f_pi"], rp["f_pi"]) < 1e-8, (t, rel_l2(gp["f_pi"], rp["f_pi"]))
        assert rel_l2(gp["f_pp"], rp["f_pp"]) < 1e-8, (t, rel_l2(gp["f_pp"], rp["f_pp"]))


def test_crowded_spread_lines_32(gpu):
    """8,000 small spheres (r = 0.5) in a 32^3 box: about 31 spread entries
    per grid line on average, so that many lines exceed the line table's 32
    slots and take the exact overflow path of the spread, and the spheres in
    the overflow lists get no corner copies from the force pass (stage 1
    gathers their corner values from the fields).  Contact cells of 2 dx
    hold two spheres on average, so the contact search stays on its fast
    path.  FP64 against the oracle."""
    print("crowded spread lines 32^3 + 8000", flush=True)
    n, npart = 32, 8000
    p, phi0, pos = make_case(n, npart, r_p=0.5, steps=20, seed=43, gap=-0.3, cell=2.0)
    for t, g_phi, r_phi, g_u, r_u, gp, rp in _run_both(p, phi0, pos, 6, chunks=3):
        assert rel_l2(g_phi, r_phi) < TOL, (t, rel_l2(g_phi, r_phi))
        assert rel_l2(g_u, r_u) < TOL, (t, rel_l2(g_u, r_u))
        assert rel_l2(gp["p"], rp["p"]) < TOL, (t, rel_l2(gp["p"], rp["p"]))
        assert rel_l2(gp["f_pi"], rp["f_pi"]) < 1e-8, (t, rel_l2(gp["f_pi"], rp["f_pi"]))
        assert rel_l2(gp["f_pp"], rp["f_pp"]) < 1e-8, (t, rel_l2(gp["f_pp"], rp["f_pp"]))


def test_heavily_crowded_spread_lines_32(gpu):
    """16,000 spheres of radius 0.5 in 32^3 (2-dx contact cells): 62 spread
    entries per grid line on average and at least 43 on every line, so every
    line's spread takes the exact path over its overflow chain (31,000
    entries in the overflow list).  FP64 against the oracle every two
    steps."""
    print("heavily crowded spread lines 32^3 + 16000", flush=True)
    n, npart = 32, 16000
    p, phi0, pos = make_case(n, npart, r_p=0.5, steps=20, seed=53, gap=-0.3, cell=2.0)
    for t, g_phi, r_phi, g_u, r_u, gp, rp in _run_both(p, phi0, pos, 4, chunks=2):
        print(f"t={t}", flush=True)
        assert rel_l2(g_phi, r_phi) < TOL, (t, rel_l2(g_phi, r_phi))
        assert rel_l2(g_u, r_u) < TOL, (t, rel_l2(g_u, r_u))
        assert rel_l2(gp["p"], rp["p"]) < TOL, (t, rel_l2(gp["p"], rp["p"]))
        assert rel_l2(gp["f_pi"], rp["f_pi"]) < 1e-8, (t, rel_l2(gp["f_pi"], rp["f_pi"]))
        assert rel_l2(gp["f_pp"], rp["f_pp"]) < 1e-8, (t, rel_l2(gp["f_pp"], rp["f_pp"]))

