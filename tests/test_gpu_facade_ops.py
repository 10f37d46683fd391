"""The facade's device field algebra (fipi_field_op, fipi_field_fft) against
NumPy restatements of the reference's operators (src/field.cpp, cited per
case): complex products when an operand is in K space, imaginary parts set
to zero by real-space operators, division by zero giving zero, products
of dim1 x dim2 components stored at (d1+1)(d2+1)-1, and the complex 3-D
DFT of each component (FFTW_FORWARD unnormalised, FFTW_BACKWARD divided by
G).  Operands are random complex fields with non-zero imaginary parts, so
the real-space rules are visible."""
import ctypes as C

import numpy as np
import pytest

import fipi_amd
from fipi_amd import config_params

pytestmark = pytest.mark.gpu
OP = {name: i for i, name in enumerate(
    ["ADD", "SUB", "ADD_S", "SUB_S", "MUL_S", "DIV_S", "MULEQ", "DIVEQ", "MUL", "DIV", "DOT", "POWER",
     "GRAD", "COPY", "RESET", "REAL", "FROM_REAL"])}
SHAPES = [(16, 16, 16), (48, 80, 96)]


class Ctx:
    def __init__(self, shape, gpu):
        import torch
        nx, ny, nz = shape
        p = config_params(nx, 0, 2.0, 10).replace(ny=ny, nz=nz, ly=float(ny), lz=float(nz))
        self.p, self.sim, self.torch, self.gpu = p, fipi_amd.Simulation(p), torch, gpu
        self.G = nx * ny * nz
        self.shape = (nz, ny, nx)

    def dev(self, z):
        """complex (dim, G) -> device float64 (dim, G, 2)"""
        a = np.stack([z.real, z.imag], -1)
        return self.torch.tensor(np.ascontiguousarray(a), dtype=self.torch.float64, device=self.gpu)

    def host(self, t):
        a = t.cpu().numpy()
        return a[..., 0] + 1j * a[..., 1]

    def op(self, code, out, out_dim, a=None, a_dim=0, a_k=0, b=None, b_dim=0, b_k=0, s=0.0):
        ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
        self.sim._c(fipi_amd.lib().fipi_field_op(self.sim._h, OP[code], ptr(out), out_dim, ptr(a), a_dim, a_k,
                                                 ptr(b), b_dim, b_k, s))

    def close(self):
        self.sim.close()


def _rand(rng, dim, G, zeros=0):
    z = rng.standard_normal((dim, G)) + 1j * rng.standard_normal((dim, G))
    if zeros:
        z.real[:, rng.choice(G, size=zeros, replace=False)] = 0.0
    return z


def _close(got, want, tol=1e-14):
    scale = max(np.abs(want).max(), 1e-300)
    assert np.abs(got - want).max() <= tol * scale, np.abs(got - want).max() / scale


def _realify(z):
    return z.real + 0j


@pytest.mark.parametrize("shape", SHAPES)
def test_in_place_operators(gpu, shape):
    """+=, -= (field.cpp:525-640), += / -= double (:574-662), *= / /=
    double (:711-830), *= / /= scalar field (:666-786), in both spaces."""
    print(f"facade in-place operators {shape}", flush=True)
    ctx = Ctx(shape, gpu)
    G = ctx.G
    rng = np.random.default_rng(1)
    for dim in (1, 3):
        x, y = _rand(rng, dim, G), _rand(rng, dim, G)
        sc = _rand(rng, 1, G, zeros=7)
        for k in (0, 1):
            cases = [("ADD", y, dim, (x + y) if k else _realify(x + y)),
                     ("SUB", y, dim, (x - y) if k else _realify(x - y)),
                     ("MULEQ", sc, 1, x * sc),  # complex when either is in K space
                     ("DIVEQ", sc, 1, None)]
            for code, b, bd, want in cases:
                t = ctx.dev(x)
                bk = k if code in ("ADD", "SUB") else (1 if code == "MULEQ" else 0)
                ctx.op(code, t, dim, a_k=k, b=ctx.dev(b), b_dim=bd, b_k=bk)
                if code == "DIVEQ":
                    r = sc.real[0]
                    safe = np.where(r != 0.0, r, 1.0)
                    want = np.where(r != 0.0, x / safe if k else (x.real / safe) + 0j, 0.0)
                _close(ctx.host(t), want)
            if not k:  # MULEQ of two real-space fields: real product, imaginary part zero
                t = ctx.dev(x)
                ctx.op("MULEQ", t, dim, a_k=0, b=ctx.dev(sc), b_dim=1, b_k=0)
                _close(ctx.host(t), _realify(x.real * sc.real + 0j))
            for code, s, want in [("ADD_S", 0.75, x.real + 0.75), ("SUB_S", 0.75, x.real - 0.75),
                                  ("MUL_S", -1.5, x * -1.5 if k else _realify(x * -1.5)),
                                  ("DIV_S", 4.0, x / 4.0 if k else _realify(x / 4.0)),
                                  ("DIV_S", 0.0, np.zeros_like(x))]:
                if k and code in ("ADD_S", "SUB_S"):
                    continue  # refused in K space (field.cpp:579-583)
                t = ctx.dev(x)
                ctx.op(code, t, dim, a_k=k, s=s)
                _close(ctx.host(t), want + 0j)
    # the space rules the reference enforces with exit(0) are errors here
    t = ctx.dev(_rand(rng, 1, G))
    with pytest.raises(fipi_amd.FipiError):
        ctx.op("ADD", t, 1, a_k=0, b=ctx.dev(_rand(rng, 1, G)), b_dim=1, b_k=1)
    with pytest.raises(fipi_amd.FipiError):
        ctx.op("ADD_S", t, 1, a_k=1, s=1.0)
    with pytest.raises(fipi_amd.FipiError):
        ctx.op("DIVEQ", t, 1, a_k=0, b=ctx.dev(_rand(rng, 1, G)), b_dim=1, b_k=1)
    ctx.close()


@pytest.mark.parametrize("shape", SHAPES)
def test_products_dot_power_grad(gpu, shape):
    """field * field and field / field (field.cpp:946-1046), dot (:364-405),
    power (:408-431), grad with the wave-vector field (:288-328), copy,
    reset and the real-part transfers."""
    print(f"facade products {shape}", flush=True)
    ctx = Ctx(shape, gpu)
    G, p = ctx.G, ctx.p
    rng = np.random.default_rng(2)
    for (da, db) in ((1, 1), (3, 1), (1, 3), (3, 3)):
        a, b = _rand(rng, da, G), _rand(rng, db, G, zeros=5)
        for ak, bk in ((0, 0), (1, 0), (0, 1), (1, 1)):
            out = ctx.dev(np.zeros((da * db, G), complex))
            ctx.op("MUL", out, da * db, a=ctx.dev(a), a_dim=da, a_k=ak, b=ctx.dev(b), b_dim=db, b_k=bk)
            want = np.zeros((da * db, G), complex)
            for d1 in range(da):
                for d2 in range(db):
                    want[(d1 + 1) * (d2 + 1) - 1] = a[d1] * b[d2] if (ak or bk) else a[d1].real * b[d2].real
            _close(ctx.host(out), want)
            if bk == 0:
                out = ctx.dev(np.zeros((da * db, G), complex))
                ctx.op("DIV", out, da * db, a=ctx.dev(a), a_dim=da, a_k=ak, b=ctx.dev(b), b_dim=db, b_k=0)
                want = np.zeros((da * db, G), complex)
                for d1 in range(da):
                    for d2 in range(db):
                        r = b[d2].real
                        safe = np.where(r != 0.0, r, 1.0)
                        q = a[d1] / safe if ak else a[d1].real / safe + 0j
                        want[(d1 + 1) * (d2 + 1) - 1] = np.where(r != 0.0, q, 0.0)
                _close(ctx.host(out), want)
        if da == db:
            for ak, bk in ((0, 0), (1, 0), (1, 1)):
                out = ctx.dev(np.zeros((1, G), complex))
                ctx.op("DOT", out, 1, a=ctx.dev(a), a_dim=da, a_k=ak, b=ctx.dev(b), b_dim=db, b_k=bk)
                want = (a * b).sum(0) if (ak or bk) else (a.real * b.real).sum(0) + 0j
                _close(ctx.host(out)[0], want)
    x = _rand(rng, 3, G)
    out = ctx.dev(np.zeros((1, G), complex))
    ctx.op("POWER", out, 1, a=ctx.dev(x), a_dim=3, a_k=0, s=3.0)
    _close(ctx.host(out)[0], (x.real ** 3).sum(0) + 0j, 1e-13)
    # grad: the reference's wave numbers, zero on each axis' Nyquist plane
    nz, ny, nx = ctx.shape
    kz, ky, kx = np.meshgrid(*(np.where(np.arange(N) <= N // 2, np.arange(N), np.arange(N) - N) * 2 * p.pi / L
                               for N, L in ((nz, p.lz), (ny, p.ly), (nx, p.lx))), indexing="ij")
    kf = np.stack([kx.reshape(-1), ky.reshape(-1), kz.reshape(-1)]) + 0j
    c = _rand(rng, 1, G)
    out = ctx.dev(np.zeros((3, G), complex))
    ctx.op("GRAD", out, 3, a=ctx.dev(c), a_dim=1, a_k=1, b=ctx.dev(kf), b_dim=3, b_k=0)
    want = 1j * kf.real * c
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for d, idx, N in ((0, ix, nx), (1, iy, ny), (2, iz, nz)):
        want[d, (idx == N // 2).reshape(-1)] = 0.0
    _close(ctx.host(out), want)
    # copy, reset, real parts out and in
    t = ctx.dev(np.zeros((3, G), complex))
    ctx.op("COPY", t, 3, a=ctx.dev(x), a_dim=3)
    assert np.array_equal(ctx.host(t), x)
    ctx.op("RESET", t, 3)
    assert not ctx.host(t).any()
    r = ctx.torch.zeros((3, G), dtype=ctx.torch.float64, device=gpu)
    ctx.op("REAL", r, 3, a=ctx.dev(x), a_dim=3)
    assert np.array_equal(r.cpu().numpy(), x.real)
    ctx.op("FROM_REAL", t, 3, a=r, a_dim=3)
    assert np.array_equal(ctx.host(t), x.real + 0j)
    ctx.close()


@pytest.mark.parametrize("shape", SHAPES)
def test_complex_transform(gpu, shape):
    """field::fft / field::ifft (field.cpp:438-455, primitive.cpp:54-62) of
    complex components with non-zero imaginary parts, against numpy.fft."""
    print(f"facade complex fft {shape}", flush=True)
    ctx = Ctx(shape, gpu)
    G = ctx.G
    rng = np.random.default_rng(3)
    x = _rand(rng, 3, G)
    t = ctx.dev(x)
    ctx.sim._c(fipi_amd.lib().fipi_field_fft(ctx.sim._h, C.c_void_p(t.data_ptr()), 3, -1))
    X = ctx.host(t)
    want = np.stack([np.fft.fftn(x[d].reshape(ctx.shape)).reshape(-1) for d in range(3)])
    assert np.linalg.norm(X - want) / np.linalg.norm(want) < 1e-14
    ctx.sim._c(fipi_amd.lib().fipi_field_fft(ctx.sim._h, C.c_void_p(t.data_ptr()), 3, 1))
    y = ctx.host(t)
    assert np.linalg.norm(y - x) / np.linalg.norm(x) < 1e-14
    # a spectrum that is not Hermitian: the inverse keeps its imaginary part
    Z = _rand(rng, 1, G)
    t = ctx.dev(Z)
    ctx.sim._c(fipi_amd.lib().fipi_field_fft(ctx.sim._h, C.c_void_p(t.data_ptr()), 1, 1))
    want = np.fft.ifftn(Z[0].reshape(ctx.shape)).reshape(-1)
    assert np.linalg.norm(ctx.host(t)[0] - want) / np.linalg.norm(want) < 1e-14
    ctx.close()
