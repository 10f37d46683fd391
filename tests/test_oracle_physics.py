"""Pins of the oracle's restatement of the physics the reference carries but
main() never calls (CPU only):

* field::interpolate2 (field.cpp:251-281): a constant field comes back
  exactly (the weights sum to one) and an affine field gives the closed
  form of the reference's stencil -- nodes floor(x)+1 and floor(x)+2 (the
  reference subtracts (1-size)/2 = -1) with fractions m/3 + 1/3;
* element::get_f_rand (element.cpp:109-144): every drawn g is one of the
  values sqrt(-2 log r1) cos(2 pi r2), r1, r2 in {0, .., 999}/1000, and
  the draws follow the exact distribution of g given by enumerating all
  10^6 pairs (r1 = 0 rejected), in a chi-square test over equal-mass bins,
  with sample moments inside their standard errors and no correlation
  between neighbouring draws;
* in the step: uf is u interpolated (element.cpp:43-49) at the positions
  the stage starts from, f_rand is the draw of the step index and Heun
  stage, and f_drag = -6 pi r mu (u - uf) (element.cpp:147-151) equals
  -(f_pi + f_rand + f_pp), because the stage velocity is
  uf + (f_pi + f_ext + f_pp)/(6 pi mu r) (element.cpp:156-173);
* with interp_u = 0 and kt = 0 the step is the reference step bit for bit.
"""
import numpy as np
import pytest

import oracle
from _cases import make_case
from fipi_amd import config_params
from fipi_amd.params import REFERENCE_PI


def test_interp2_constant_and_affine_closed_form():
    p = config_params(16, 0, 2.0, 10).replace(ny=12, nz=20, ly=12.0, lz=20.0)
    nx, ny, nz = p.nx, p.ny, p.nz
    rng = np.random.default_rng(3)
    # constants: the eight weights sum to one
    pos = rng.uniform(0, 1, (500, 3)) * [p.lx, p.ly, p.lz]
    const = np.full((nz, ny, nx), 1.75)
    assert np.abs(oracle.interp2(p, const, pos) - 1.75).max() < 1e-14
    # affine: value(i, j, k) = a i + b j + c k on the nodes; positions kept
    # away from the last two nodes of each axis so that no index wraps
    a, b, c = 0.7, -1.3, 2.1
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    grid = a * i + b * j + c * k
    pos = rng.uniform(0, 1, (2000, 3)) * [nx - 3, ny - 3, nz - 3]
    fl = np.floor(pos)
    frac = (pos - fl) / 3 + 1.0 / 3
    want = ((fl + 1 + frac) * [a, b, c]).sum(1)
    got = oracle.interp2(p, grid, pos)
    assert np.abs(got - want).max() < 1e-12
    # field::interpolate (nodes floor, floor+1) reproduces the affine field;
    # interpolate2 is shifted along each axis by 4/3 - 2m/3 cells
    lin = oracle.interp(p, grid, pos)
    assert np.abs(lin - (pos * [a, b, c]).sum(1)).max() < 1e-12
    shift = 4.0 / 3 - 2.0 * (pos - fl) / 3
    assert np.abs(got - lin - (shift * [a, b, c]).sum(1)).max() < 1e-12
    # periodic wrap: the nodes of a position in the last cell wrap to 0, 1
    q = np.array([[nx - 0.5, 0.5, 0.5]])
    g1 = np.zeros((nz, ny, nx))
    g1[1, 1, 0] = 1.0  # node i = (nx-1)+1 = 0, j = 1, k = 1
    m = 0.5 / 3 + 1.0 / 3
    assert abs(oracle.interp2(p, g1, q)[0] - (1 - m) ** 3) < 1e-15


def _all_g(pi=REFERENCE_PI):
    r = np.arange(1000) / 1000
    r1, r2 = np.meshgrid(r, r, indexing="ij")
    with np.errstate(divide="ignore", invalid="ignore"):
        g = np.sqrt(-2 * np.log(r1)) * np.cos(2 * pi * r2)
    keep = ~(np.abs(g) > 5.0)
    return np.sort(g[keep])


def test_brownian_draws_follow_the_reference_distribution():
    p = config_params(16, 0, 2.0, 10).replace(kt=0.25, noise_seed=12345)
    amp = np.sqrt(12.0 * p.pi * p.r_p * p.mu * p.kt / p.dt)
    n = 20000
    draws = np.concatenate([oracle.brownian(p, t, s, n).reshape(-1) for t in (0, 1, 7) for s in (1, 2)])
    g = draws / amp
    allg = _all_g()
    assert allg.size == 999000  # r1 = 0 is the only rejection
    # every draw is one of the reference's values
    idx = np.clip(np.searchsorted(allg, g), 1, allg.size - 1)
    near = np.minimum(np.abs(allg[idx] - g), np.abs(allg[idx - 1] - g))
    assert near.max() < 1e-12
    # chi-square over 100 bins of equal probability under the exact law
    nb = 100
    edges = allg[(np.arange(1, nb) * allg.size) // nb]
    counts = np.bincount(np.searchsorted(edges, g, side="right"), minlength=nb)
    exact = np.diff(np.concatenate([[0], np.searchsorted(allg, edges, side="right"), [allg.size]])) / allg.size
    chi2 = ((counts - exact * g.size) ** 2 / (exact * g.size)).sum()
    assert chi2 < 160, chi2  # 99 degrees of freedom: p ~ 1e-4
    # moments against the enumerated ones
    m1, m2, m4 = allg.mean(), (allg ** 2).mean(), (allg ** 4).mean()
    N = g.size
    assert abs(g.mean() - m1) < 5 * np.sqrt(m2 / N)
    assert abs((g ** 2).mean() - m2) < 5 * np.sqrt((m4 - m2 ** 2) / N)
    # neighbouring draws (components, particles) are uncorrelated
    assert abs(np.corrcoef(g[:-1], g[1:])[0, 1]) < 5 / np.sqrt(N)
    # another seed, step or stage gives other draws
    assert not np.array_equal(oracle.brownian(p, 0, 1, 50), oracle.brownian(p, 0, 2, 50))
    assert not np.array_equal(oracle.brownian(p, 0, 1, 50), oracle.brownian(p, 1, 1, 50))
    assert not np.array_equal(oracle.brownian(p, 0, 1, 50), oracle.brownian(p, 0, 1, 50, noise_seed=1))
    assert not oracle.brownian(p, 0, 1, 50, kt=0.0).any()


def _wrap(x, L):
    return np.fmod(x + L, L)


@pytest.mark.parametrize("mode", [1, 2])
def test_step_uf_brownian_and_drag(mode):
    p, phi0, pos = make_case(32, 300, steps=20, seed=51, gap=-0.5, amp=0.3)
    p = p.replace(interp_u=mode, kt=0.05, noise_seed=99)
    s = oracle.Sim(p, phi0, pos)
    steps = 3
    s.step(steps)
    part = s.particles()
    fl = s.particle_fluid()
    u = s.u()
    L = np.array([p.lx, p.ly, p.lz])
    # stage 2 starts from the stage-1 positions p_old + dt u_old, wrapped
    p1 = _wrap(part["p_old"] + p.dt * part["u_old"], L)
    ip = oracle.interp if mode == 1 else oracle.interp2
    uf_want = np.stack([ip(p, u[d], p1) for d in range(3)], 1)
    assert np.abs(fl["uf"] - uf_want).max() <= 1e-14 * np.abs(uf_want).max()
    assert np.abs(fl["uf"]).max() > 0
    # the draws of the last step's second stage
    assert np.array_equal(fl["f_rand"], oracle.brownian(p, steps - 1, 2, p.n_particles))
    # drag from the stage velocity balances the other forces
    drag = oracle.drag(p, part["u"], fl["uf"])
    F = part["f_pi"] + fl["f_rand"] + part["f_pp"]
    assert np.abs(drag + F).max() <= 1e-12 * np.abs(F).max()
    assert np.abs(part["f_pp"]).max() > 0 and np.abs(part["f_pi"]).max() > 0
    # the drift uf moves the particles: against a run without it
    s0 = oracle.Sim(p.replace(interp_u=0), phi0, pos)
    s0.step(steps)
    assert np.abs(s0.particles()["p"] - part["p"]).max() > 1e-6


def test_defaults_are_the_reference_step():
    p, phi0, pos = make_case(32, 200, steps=20, seed=52, gap=-0.5)
    a = oracle.Sim(p, phi0, pos)
    b = oracle.Sim(p.replace(interp_u=0, kt=0.0, noise_seed=77), phi0, pos)
    a.step(4)
    b.step(4)
    assert np.array_equal(a.phi(), b.phi())
    assert np.array_equal(a.particles()["p"], b.particles()["p"])
    fl = a.particle_fluid()
    assert not fl["uf"].any() and not fl["f_rand"].any()
    # kt only: the Brownian force alone moves the particles
    c = oracle.Sim(p.replace(kt=0.05, noise_seed=77), phi0, pos)
    c.step(4)
    assert np.abs(c.particles()["p"] - a.particles()["p"]).max() > 1e-6
