"""The physics the reference carries but main() never calls, on the GPU
against the oracle: field::interpolate2 (field.cpp:251-281), the Brownian
force of element::get_f_rand (element.cpp:109-144) with its counter-based
draws, and the step with uf interpolated from u (element.cpp:43-49, either
interpolation) and the Brownian force in the f_ext slot of the advection,
with f_drag = -6 pi r mu (u - uf) (element.cpp:147-151) read back."""
import numpy as np
import pytest

import fipi_amd
import oracle
from _cases import make_case, rel_l2
from fipi_amd import FP32_MIXED, FP64, config_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [FP64, FP32_MIXED])
def test_interp2_unit(gpu, prec):
    print(f"interp2 unit prec={prec}", flush=True)
    p = config_params(32, 0, 2.0, 10, precision=prec).replace(ny=48, ly=48.0)
    rng = np.random.default_rng(61)
    grid = rng.standard_normal((p.nz, p.ny, p.nx))
    pos = rng.uniform(0, 1, (5000, 3)) * [p.lx, p.ly, p.lz]
    pos[:8] = [[p.lx - 1e-9, 0.0, 5.0], [0.0, p.ly - 0.5, p.lz - 0.25]] * 4  # wrapped nodes
    sim = fipi_amd.Simulation(p)
    got = sim.interp2(grid, pos)
    sim.close()
    ref = oracle.interp2(p, grid.astype(np.float32).astype(np.float64) if prec else grid, pos)
    assert np.abs(got - ref).max() <= (1e-14 if prec == FP64 else 1e-12) * np.abs(ref).max()


def test_brownian_unit(gpu):
    print("brownian unit", flush=True)
    p = config_params(32, 0, 2.0, 10).replace(kt=0.3, noise_seed=2024)
    sim = fipi_amd.Simulation(p)
    for t, stage in ((0, 1), (0, 2), (123, 1)):
        got = sim.brownian(t, stage, 30000)
        ref = oracle.brownian(p, t, stage, 30000)
        # same integer draws; log / cos / sqrt of the device and of glibc
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    sim.close()


@pytest.mark.parametrize("mode,kt", [(1, 0.0), (2, 0.0), (0, 0.05), (1, 0.05), (2, 0.05)])
def test_step_with_uf_and_noise_vs_oracle(gpu, mode, kt):
    print(f"step interp_u={mode} kt={kt}", flush=True)
    p, phi0, pos = make_case(32, 300, steps=40, seed=62, gap=-0.5, amp=0.3)
    p = p.replace(interp_u=mode, kt=kt, noise_seed=7)
    cth = np.cos(np.deg2rad(fipi_amd.init_uniform(300, 60.0, 120.0, 63)))
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0, pos, None, cth)
    sim.step(12)
    ref = oracle.Sim(p, phi0, pos, None, cth)
    ref.step(12)
    assert rel_l2(sim.phi(), ref.phi()) < 1e-10
    assert rel_l2(sim.velocity(), ref.u()) < 1e-10
    gp, rp = sim.particles(), ref.particles()
    assert rel_l2(gp["p"], rp["p"]) < 1e-10
    assert rel_l2(gp["u"], rp["u"]) < 1e-10
    g, r = sim.particle_fluid(), ref.particle_fluid()
    if mode:
        assert rel_l2(g["uf"], r["uf"]) < 1e-10
    else:
        assert not g["uf"].any()
    if kt:
        assert rel_l2(g["f_rand"], r["f_rand"]) < 1e-12
    else:
        assert not g["f_rand"].any()
    drag = oracle.drag(p, rp["u"], r["uf"])
    assert np.abs(g["f_drag"] - drag).max() <= 1e-9 * np.abs(drag).max()
    sim.close()


def test_mixed_with_uf_and_noise_vs_oracle(gpu):
    print("mixed step interp_u=1 kt=0.05", flush=True)
    p, phi0, pos = make_case(32, 300, steps=40, seed=64, gap=-0.5, amp=0.3)
    p = p.replace(interp_u=1, kt=0.05, noise_seed=8)
    sim = fipi_amd.Simulation(p.replace(precision=FP32_MIXED))
    sim.set_state(phi0, pos)
    sim.step(12)
    ref = oracle.Sim(p, phi0, pos)
    ref.step(12)
    assert rel_l2(sim.phi(), ref.phi()) < 1e-5
    assert np.abs(sim.particles()["p"] - ref.particles()["p"]).max() < 1e-4
    sim.close()


def test_batched_paths_refuse_the_extra_physics(gpu):
    common = config_params(32, 0, 2.0, 10, precision=FP32_MIXED).replace(switch_particle=1, interp_u=1)
    with pytest.raises(fipi_amd.FipiError):
        fipi_amd.Ensemble(common, [32, 32], [10, 10])
    p, phi0, pos = make_case(32, 50, steps=20, seed=65)
    with pytest.raises(fipi_amd.FipiError):
        fipi_amd.Adjoint(p.replace(kt=0.1), phi0, pos, 0.5, 10, 5)
