"""Hand-written adjoint of the CH/Stokes step (fipi_adjoint_*) for the
calibration of mobility, lambda and per-particle contact angles.

Checks: central finite differences (h = 1e-4) of the FP64 forward run, which
itself matches the oracle step (test_gpu_step.py); bitwise independence of
the checkpoint interval (every segment is recomputed exactly); the
FP32-mixed adjoint against the FP64 one; the torch.autograd.Function
harness."""
import numpy as np
import pytest

import fipi_amd
from _cases import rel_l2
from fipi_amd import FP32_MIXED, FP64, config_params

pytestmark = pytest.mark.gpu

A_W = 0.5


def _problem(prec, n=32, npart=40, steps=30, every=7, seed=4):
    p = config_params(n, npart, 2.0, steps, precision=prec)
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, seed)
    pos = fipi_amd.init_rsa(npart, 2.0, 0.0, (n, n, n), seed + 1)
    assert pos.shape[0] == npart
    th_true = np.deg2rad(fipi_amd.init_uniform(npart, 70.0, 110.0, seed + 2))
    adj = fipi_amd.Adjoint(p, phi0, pos, A_W, steps, every)
    _, target = adj.forward(1.3, 0.8, th_true, want_loss=False)
    adj.set_target(target)
    th0 = np.deg2rad(90.0 + 5.0 * np.sin(np.arange(npart)))
    return adj, th0


def _fd(adj, m, lam, th, which, h=1e-4, idx=None):
    def loss(mm, ll, tt):
        return adj.forward(mm, ll, tt)[0]
    if which == "m":
        return (loss(m + h, lam, th) - loss(m - h, lam, th)) / (2 * h)
    if which == "lam":
        return (loss(m, lam + h, th) - loss(m, lam - h, th)) / (2 * h)
    tp, tm = th.copy(), th.copy()
    tp[idx] += h
    tm[idx] -= h
    return (loss(m, lam, tp) - loss(m, lam, tm)) / (2 * h)


def test_adjoint_vs_central_differences_fp64(gpu):
    print("adjoint FD fp64", flush=True)
    adj, th0 = _problem(FP64)
    m, lam = 1.0, 1.0
    loss, gm, gl, gth = adj.gradient(m, lam, th0)
    assert loss > 0
    l2 = adj.forward(m, lam, th0)[0]
    assert abs(l2 - loss) <= 1e-12 * loss
    fm = _fd(adj, m, lam, th0, "m")
    fl = _fd(adj, m, lam, th0, "lam")
    assert abs(gm - fm) <= 1e-6 * abs(fm), (gm, fm)
    assert abs(gl - fl) <= 1e-6 * abs(fl), (gl, fl)
    # the particles with the largest contact-angle sensitivity
    big = np.argsort(-np.abs(gth))[:4]
    gscale = np.abs(gth).max()
    for i in big:
        ft = _fd(adj, m, lam, th0, "th", idx=int(i))
        assert abs(gth[i] - ft) <= 1e-6 * gscale, (i, gth[i], ft)
    adj.close()


def test_checkpoint_interval_is_exact(gpu):
    print("adjoint checkpoints", flush=True)
    a1, th0 = _problem(FP64, steps=24, every=5)
    a2, _ = _problem(FP64, steps=24, every=24)
    r1 = a1.gradient(1.0, 1.0, th0)
    r2 = a2.gradient(1.0, 1.0, th0)
    assert r1[:3] == r2[:3]
    assert np.array_equal(r1[3], r2[3])
    a1.close()
    a2.close()


def test_first_step_only(gpu):
    """T = 1: only the first-order step (no history, no convection)."""
    print("adjoint T=1", flush=True)
    adj, th0 = _problem(FP64, steps=1, every=1)
    loss, gm, gl, gth = adj.gradient(1.0, 1.0, th0)
    fm = _fd(adj, 1.0, 1.0, th0, "m")
    fl = _fd(adj, 1.0, 1.0, th0, "lam")
    assert abs(gm - fm) <= 1e-6 * abs(fm), (gm, fm)
    assert abs(gl - fl) <= 1e-6 * abs(fl), (gl, fl)
    adj.close()


def test_mixed_vs_fp64_gradient(gpu):
    print("adjoint mixed", flush=True)
    a64, th0 = _problem(FP64)
    a32, _ = _problem(FP32_MIXED)
    l64, m64, k64, t64 = a64.gradient(1.0, 1.0, th0)
    l32, m32, k32, t32 = a32.gradient(1.0, 1.0, th0)
    assert abs(l32 - l64) <= 1e-4 * l64
    assert abs(m32 - m64) <= 1e-3 * abs(m64)
    assert abs(k32 - k64) <= 1e-3 * abs(k64)
    assert np.linalg.norm(t32 - t64) <= 1e-3 * np.linalg.norm(t64)
    a64.close()
    a32.close()


def test_torch_autograd_harness(gpu):
    print("adjoint torch", flush=True)
    import torch
    from fipi_amd.autograd import calibration_loss
    adj, th0 = _problem(FP64, steps=10, every=4)
    m = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    lam = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    th = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    loss = calibration_loss(m, lam, th, adj)
    loss.backward()
    l0, gm, gl, gth = adj.gradient(1.0, 1.0, th0)
    assert float(loss) == l0
    assert float(m.grad) == gm and float(lam.grad) == gl
    assert np.array_equal(th.grad.numpy(), gth)
    # parameters on the GPU, theta in float32: gradients come back on the
    # inputs' devices and dtypes
    m = torch.tensor(1.0, dtype=torch.float64, device="cuda", requires_grad=True)
    lam = torch.tensor(1.0, dtype=torch.float64, device="cuda", requires_grad=True)
    th = torch.tensor(th0, dtype=torch.float32, device="cuda", requires_grad=True)
    loss = calibration_loss(m, lam, th, adj)
    loss.backward()
    assert loss.device.type == "cuda"
    assert m.grad.device.type == "cuda" and th.grad.device.type == "cuda" and th.grad.dtype == torch.float32
    l1, gm1, gl1, gth1 = adj.gradient(1.0, 1.0, th.detach().cpu().double().numpy())
    assert float(m.grad) == gm1 and float(lam.grad) == gl1
    assert np.allclose(th.grad.cpu().double().numpy(), gth1, rtol=1e-6, atol=1e-6 * np.abs(gth1).max())
    adj.close()


def _oracle_loss(p, phi0, pos, theta, m, lam, target, steps):
    import oracle
    q = p.replace(mobility=m, lam=lam)
    s = oracle.Sim(q, phi0, pos, None, np.cos(theta), wetting=A_W, fix_particles=True)
    s.step(steps)
    phi = s.phi()
    return float(((phi - target) ** 2).sum()), phi


def test_adjoint_forward_and_gradient_vs_oracle(gpu):
    """The adjoint's forward model (wetting source a_w = 0.5, particles held
    fixed with full-ramp contact forces) against the oracle's calibration
    model, and the adjoint gradient against FP64 central differences of
    the oracle itself (the FP64 reference gradient)."""
    print("adjoint vs oracle calibration model", flush=True)
    n, npart, steps, seed = 32, 40, 30, 4
    p = config_params(n, npart, 2.0, steps, precision=FP64)
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, seed)
    pos = fipi_amd.init_rsa(npart, 2.0, -0.5, (n, n, n), seed + 1)
    assert pos.shape[0] == npart
    th_true = np.deg2rad(fipi_amd.init_uniform(npart, 70.0, 110.0, seed + 2))
    _, target = _oracle_loss(p, phi0, pos, th_true, 1.3, 0.8, np.zeros((n, n, n)), steps)
    adj = fipi_amd.Adjoint(p, phi0, pos, A_W, steps, 7)
    adj.set_target(target)
    th0 = np.deg2rad(90.0 + 5.0 * np.sin(np.arange(npart)))
    # forward at two parameter sets against the oracle
    for m, lam, th in ((1.0, 1.0, th0), (1.3, 0.8, th_true)):
        g_loss, g_phi = adj.forward(m, lam, th)
        r_loss, r_phi = _oracle_loss(p, phi0, pos, th, m, lam, target, steps)
        assert rel_l2(g_phi, r_phi) < 1e-10, (m, lam, rel_l2(g_phi, r_phi))
        if r_loss > 0:
            assert abs(g_loss / r_loss - 1) < 1e-8, (g_loss, r_loss)
    # the FP64 reference gradient: central differences of the oracle
    loss, gm, gl, gth = adj.gradient(1.0, 1.0, th0)
    h = 1e-4
    f = lambda mm, ll, tt: _oracle_loss(p, phi0, pos, tt, mm, ll, target, steps)[0]
    fm = (f(1.0 + h, 1.0, th0) - f(1.0 - h, 1.0, th0)) / (2 * h)
    fl = (f(1.0, 1.0 + h, th0) - f(1.0, 1.0 - h, th0)) / (2 * h)
    print(f"dM {gm:.10e} oracle fd {fm:.10e}; dlambda {gl:.10e} oracle fd {fl:.10e}", flush=True)
    assert abs(gm - fm) <= 1e-6 * abs(fm)
    assert abs(gl - fl) <= 1e-6 * abs(fl)
    gs = np.abs(gth).max()
    for i in np.argsort(-np.abs(gth))[:3]:
        tp, tm = th0.copy(), th0.copy()
        tp[i] += h
        tm[i] -= h
        fd = (f(1.0, 1.0, tp) - f(1.0, 1.0, tm)) / (2 * h)
        assert abs(gth[i] - fd) <= 1e-6 * gs, (i, gth[i], fd)
    adj.close()
