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
