"""Closed-form pins of the oracle's time-step composition (CPU only).

The oracle restates src/main.cpp:75-222 statement by statement.  These
tests check the parts whose order and staleness matter against answers
computed here from the mathematics, independently of the oracle's loops:

* the two Heun stages and the periodic wrap of element.cpp:156-186 (two
  isolated sphere pairs, one crossing x = 0 and one crossing x = L, in a
  uniform phi so that the attachment force vanishes), over two steps so
  that the contact ramp's t+1 argument (main.cpp:202,213) is exercised;
* the convection term of the CH update uses the velocity of the previous
  step, u^n (main.cpp:86-93), and the SBDF2 history carries conv^{n-1}
  (main.cpp:117-137): a small single Fourier mode advected by a uniform
  flow follows a scalar recurrence;
* the free energy sum lambda(|grad c|^2/2 + (c^2-1)^2/(4 eps^2)) dx^3
  (main.cpp:232, field.cpp:506-516) decreases monotonically under pure
  Cahn-Hilliard (switch_ns = 0), evaluated with a NumPy spectral gradient.
"""
import numpy as np
import pytest

import oracle
from _cases import make_case, rel_l2
from fipi_amd import config_params


def _min_image(d, L, th):
    """cluster.cpp:223-232: shift by L only beyond the threshold th."""
    d = d.copy()
    d = np.where(d > th, d - L, d)
    d = np.where(d < -th, d + L, d)
    return d


def _contact_forces_brute(p, pos, ramp):
    """All-pairs contact force of cluster.cpp:174-275 (no cell list)."""
    nc, dc = p.cell_grid()
    L = np.array([p.lx, p.ly, p.lz])
    th = np.array([dc[a] * (nc[a] - 2) for a in range(3)])
    f = np.zeros_like(pos)
    for a in range(len(pos)):
        for b in range(a + 1, len(pos)):
            d = _min_image(pos[a] - pos[b], L, th)
            nr = np.sqrt(np.sum(d * d))
            if nr < p.r_eq:
                fn = (p.r_eq - nr) * ramp * p.k_n * d / nr
                f[a] += fn
                f[b] -= fn
    return f


def _ramp(t, thr):
    # primitive.cpp:122-129
    return 0.0 if t < 0 else (t / thr if t < thr else 1.0)


def test_heun_two_pairs_closed_form():
    """Two Heun stages per step (element.cpp:156-173) with v = F/(6 pi mu r)
    (uf = 0, f_pi = 0, f_ext = 0), then x <- fmod(x + L, L)
    (element.cpp:177-186).  Sphere 2 is pushed across x = 0 and sphere 4
    across x = L; ramp_t = 4 puts the ramp at 1/4 in step 1, 1/2 in step 2."""
    n = 32
    p = config_params(n, 6, 2.0, 100).replace(switch_ns=0, ramp_t=4.0, mu=0.7, dt=0.02)
    # uniform phi = 1: |phi| > tanh 2, so f_pi = 0 (element.cpp:84); f(1) = 0
    # keeps phi at 1 under CH
    phi0 = np.ones((n, n, n))
    pos = np.array([
        [10.0, 10.3, 9.7], [12.9, 11.1, 9.2],          # generic overlap
        [3e-5, 5.0, 5.0], [2.0, 5.3, 5.2],             # 2 pushed to x < 0
        [n - 2e-5, 25.0, 25.0], [n - 1.5, 25.2, 24.9],  # 4 pushed to x > L
    ])
    s = oracle.Sim(p, phi0, pos)
    L = np.array([p.lx, p.ly, p.lz])
    r = p.r_p
    cur = pos.copy()
    for step in range(2):
        rmp = _ramp(step + 1, p.ramp_t)
        v1 = _contact_forces_brute(p, cur, rmp) / (6.0 * p.pi * p.mu * r)
        p_old = cur.copy()
        mid = np.fmod(cur + v1 * p.dt + L, L)
        v2 = _contact_forces_brute(p, mid, rmp) / (6.0 * p.pi * p.mu * r)
        new = np.fmod(p_old + 0.5 * p.dt * (v1 + v2) + L, L)
        s.step(1)
        got = s.particles()
        # displacements, minimum image: a wrong stage (p_old vs mid, u_old
        # vs u) changes them at O(1) relative
        disp_exp = new - p_old
        disp_exp -= L * np.round(disp_exp / L)
        disp_got = got["p"] - p_old
        disp_got -= L * np.round(disp_got / L)
        assert np.abs(disp_exp).max() > 1e-4
        assert rel_l2(disp_got, disp_exp) < 1e-10
        assert rel_l2(got["u_old"], v1) < 1e-12
        assert rel_l2(got["u"], v2) < 1e-12
        assert np.array_equal(got["p_old"], p_old)
        cur = new
    # the two boundary crossings happened and were wrapped into [0, L)
    assert cur[2, 0] > n - 1.0 and cur[4, 0] < 1.0
    assert (cur >= 0).all() and (cur < L).all()
    assert np.abs(s.particles()["p"] - cur).max() < 1e-12


def test_convection_uses_previous_velocity():
    """c0 = A cos(k.x + 0.7) (both parts of c_hat non-zero), u^0 = U uniform, A small (f linear in c).  Step 0:
    c_hat^1 = c_hat^0 (1 + a - i b)/(1 + dt M lam k^4) (main.cpp:106-113) with
    a = dt M k^2 lam/eps^2, b = dt k.U.  Step 1 uses u^1 from the Stokes
    solve of step 0 (O(A^2), so no convection of its own) and the history
    conv^0 (main.cpp:117-127):
    c_hat^2 = ((4 + 4a) c_hat^1 - (1 + 2a) c_hat^0 + 2 i b c_hat^0)/(3 + 2 dt M lam k^4)."""
    n = 16
    p = config_params(n, 0, 2.0, 10).replace(switch_ns=1, dt=0.05)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    m = (1, 2, 0)
    arg = 2 * np.pi * (m[0] * x + m[1] * y + m[2] * z) / n
    A = 1e-8
    ph = 0.7
    s = oracle.Sim(p, A * np.cos(arg + ph))
    U = np.array([0.3, -0.2, 0.1])
    u = np.zeros((3, n, n, n))
    for d in range(3):
        u[d] = U[d]
    s.set_u(u)
    kv = np.array(m) * 2.0 * p.pi / p.lx  # the reference's wave numbers (pi = 3.1415926)
    k2 = float(kv @ kv)
    dtM, lam, e2 = p.dt * p.mobility, p.lam, p.epsilon ** 2
    a = dtM * k2 * lam / e2
    b = p.dt * float(kv @ U)
    Z0 = A * np.exp(1j * ph)
    Z1 = Z0 * (1 + a - 1j * b) / (1 + dtM * lam * k2 * k2)
    Z2 = ((4 + 4 * a) * Z1 - (1 + 2 * a) * Z0 + 2j * b * Z0) / (3 + 2 * dtM * lam * k2 * k2)
    cb, sb = np.cos(arg), np.sin(arg)

    def amp(c):
        # c = Re(Z e^{i k.x}) = Re Z cos - Im Z sin
        return complex(np.sum(c * cb) / np.sum(cb * cb), -np.sum(c * sb) / np.sum(sb * sb))

    s.step(1)
    z1 = amp(s.phi())
    assert abs(Z1.imag) > 1e-3 * A  # the flow moved the mode
    assert abs(z1 - Z1) < 1e-10 * A
    assert np.abs(s.u()).max() < 1e-12  # u^1 = O(A^2)
    s.step(1)
    z2 = amp(s.phi())
    assert abs(z2 - Z2) < 1e-10 * A
    # a step that (wrongly) convected with U again would be this far off
    Z2_wrong = Z2 - 4j * b * Z1 / (3 + 2 * dtM * lam * k2 * k2)
    assert abs(Z2_wrong - Z2) > 1e3 * abs(z2 - Z2)


def _spectral_grad(c, p):
    """grad c with the reference's wave numbers and Nyquist zeroing
    (field.cpp:100-121, 288-328), in NumPy."""
    nz, ny, nx = c.shape
    ch = np.fft.fftn(c)
    out = []
    for axis, (N, L) in enumerate(((nx, p.lx), (ny, p.ly), (nz, p.lz))):
        idx = np.arange(N)
        k = np.where(idx <= N // 2, idx, idx - N) * 2.0 * p.pi / L
        k[N // 2] = 0.0  # the Nyquist index of each component is zeroed
        shape = [1, 1, 1]
        shape[2 - axis] = N
        out.append(np.real(np.fft.ifftn(1j * k.reshape(shape) * ch)))
    return np.array(out)


def free_energy(c, p):
    g = _spectral_grad(c, p)
    gg = (g ** 2).sum(0)
    dens = p.lam * (0.5 * gg + 0.25 / p.epsilon ** 2 * (c ** 2 - 1.0) ** 2)
    return float(dens.sum() * p.dx ** 3)


def test_free_energy_decreases_pure_ch():
    p, phi0, _ = make_case(32, 0, steps=100, ns=0, seed=9, amp=0.3)
    s = oracle.Sim(p, phi0)
    e_prev = free_energy(phi0, p)
    drops = []
    for _ in range(30):
        s.step(2)
        e = free_energy(s.phi(), p)
        assert e < e_prev
        drops.append(e_prev - e)
        e_prev = e
    # a sign error in the mobility or gradient-energy term would raise it
    assert max(drops) > 1e-3 * abs(e_prev)


@pytest.mark.parametrize("steps", [1, 3])
def test_snapshot_style_restart_of_the_oracle_is_a_pure_function_of_state(steps):
    """Two oracle runs from the same state agree bitwise (the oracle is
    deterministic, which the restart and long-horizon tests rely on)."""
    p, phi0, pos = make_case(16, 20, r_p=1.5, steps=20, ns=1, seed=4, gap=-0.5, cell=5.0)
    a = oracle.Sim(p, phi0, pos)
    b = oracle.Sim(p, phi0, pos)
    a.step(steps)
    b.step(steps)
    assert np.array_equal(a.phi(), b.phi())
    assert np.array_equal(a.particles()["p"], b.particles()["p"])


def test_calibration_model_pins():
    """The oracle's calibration model (the forward model of libfipi's
    adjoint): zero wetting weights give the reference step bitwise; the
    wetting source integrates to -a_w sum cos(theta) (the spread conserves
    the weights); with particles held fixed they do not move and their
    contact forces are the full-ramp forces of cluster.cpp:174-275."""
    p, phi0, pos = make_case(32, 60, r_p=2.0, steps=20, ns=1, seed=6, gap=-0.5)
    a = oracle.Sim(p, phi0, pos)
    b = oracle.Sim(p, phi0, pos, None, np.zeros(60), wetting=0.7)
    a.step(3)
    b.step(3)
    assert np.array_equal(a.phi(), b.phi())
    # sum W dx^3 = -a_w sum cos(theta): one step of pure diffusion-free
    # bookkeeping is not needed, the spread itself is checked
    ct = np.cos(np.deg2rad(np.linspace(60.0, 120.0, 60)))
    w = np.zeros((60, 3))
    w[:, 0] = -0.7 * ct
    W = oracle.spread(p, pos, w)[0]
    assert abs(W.sum() * p.dx ** 3 - (-0.7 * ct.sum())) < 1e-12
    # fixed particles: positions unchanged, f_pp at the full ramp, and the
    # wetting source changes phi
    c = oracle.Sim(p, phi0, pos, None, ct, wetting=0.7, fix_particles=True)
    c.step(4)
    got = c.particles()
    assert np.array_equal(got["p"], pos)
    _, f_full, _ = oracle.contact_forces(p, pos, t=int(np.ceil(p.ramp_t)) + 1)
    assert np.array_equal(got["f_pp"], f_full)
    assert np.abs(f_full).max() > 0
    d = oracle.Sim(p, phi0, pos, None, ct, wetting=0.0, fix_particles=True)
    d.step(4)
    assert rel_l2(c.phi(), d.phi()) > 1e-6
