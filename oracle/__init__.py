"""ctypes wrapper of the CPU FP64 oracle (libfipi_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``, always as the checker or the
timed CPU baseline, never as the product path.  The C sources restate the
reference time step (src/main.cpp:75-222); see oracle/fipi_oracle.h for the
per-function citations and how each part is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfipi_oracle.so")
_lib = None


class OracleParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("n_particles", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("pi", C.c_double),
        ("dt", C.c_double),
        ("ramp_t", C.c_double),
        ("epsilon", C.c_double),
        ("lambda_", C.c_double),
        ("mobility", C.c_double),
        ("mu", C.c_double),
        ("sigma", C.c_double),
        ("a_pi", C.c_double),
        ("k_n", C.c_double),
        ("r_p", C.c_double),
        ("r_eq", C.c_double),
        ("r_c", C.c_double),
        ("margin", C.c_double),
        ("rho1", C.c_double), ("rho2", C.c_double),
        ("gravity", C.c_double),
        ("tanh2", C.c_double),
        ("switch_particle", C.c_int),
        ("switch_ch", C.c_int),
        ("switch_ns", C.c_int),
        ("n_threads", C.c_int),
        ("wetting", C.c_double),
        ("fix_particles", C.c_int),
        ("interp_u", C.c_int),
        ("kt", C.c_double),
        ("noise_seed", C.c_ulonglong),
    ]


def build(force: bool = False) -> str:
    """Compile libfipi_oracle.so with oracle/Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.oracle_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int]
        L.oracle_fft3d.restype = C.c_int
        L.oracle_set_fft_order.argtypes = [C.c_int]
        L.oracle_set_fft_order.restype = None
        L.oracle_cell_grid.argtypes = [C.POINTER(OracleParams), ip, dp]
        L.oracle_cells_build.argtypes = [C.POINTER(OracleParams), C.c_int, dp, ip, ip, ip]
        L.oracle_cells_build.restype = C.c_int
        L.oracle_contact_pairs.argtypes = [C.POINTER(OracleParams), C.c_int, dp, C.c_int64, ip]
        L.oracle_contact_pairs.restype = C.c_int64
        L.oracle_contact_forces.argtypes = [C.POINTER(OracleParams), C.c_int, dp, dp,
                                            C.c_int, C.c_int, dp, dp]
        L.oracle_contact_forces.restype = C.c_int
        L.oracle_interp.argtypes = [C.POINTER(OracleParams), dp, C.c_int, dp, dp]
        L.oracle_interp2.argtypes = [C.POINTER(OracleParams), dp, C.c_int, dp, dp]
        L.oracle_brownian.argtypes = [C.POINTER(OracleParams), C.c_int, C.c_int, C.c_int, dp]
        L.oracle_drag.argtypes = [C.POINTER(OracleParams), C.c_int, dp, dp, dp, dp]
        L.oracle_get_particle_fluid.argtypes = [C.c_void_p, dp, dp]
        L.oracle_spread.argtypes = [C.POINTER(OracleParams), C.c_int, dp, dp, dp]
        L.oracle_f_pi.argtypes = [C.POINTER(OracleParams), C.c_int, dp, dp, dp, dp, dp]
        L.oracle_create.argtypes = [C.POINTER(OracleParams), dp, dp, dp, dp]
        L.oracle_create.restype = C.c_void_p
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_step.argtypes = [C.c_void_p, C.c_int]
        L.oracle_step.restype = C.c_int
        L.oracle_time.argtypes = [C.c_void_p]
        L.oracle_time.restype = C.c_int
        for nm in ("oracle_get_phi", "oracle_get_u", "oracle_get_grad_c", "oracle_set_u"):
            getattr(L, nm).argtypes = [C.c_void_p, dp]
        L.oracle_get_particles.argtypes = [C.c_void_p] + [dp] * 7
        L.oracle_get_tau.argtypes = [C.c_void_p]
        L.oracle_get_tau.restype = C.c_double
        L.oracle_stokes.argtypes = [C.POINTER(OracleParams), dp, dp]
        L.oracle_jitter.argtypes = [C.c_uint32, C.c_uint32, C.c_int, C.c_double]
        L.oracle_jitter.restype = C.c_double
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


def to_struct(p) -> OracleParams:
    """Build OracleParams from any object/dict with the same field names
    (``lambda_`` may also be given as ``lam``)."""
    get = (lambda k: p[k]) if isinstance(p, dict) else (lambda k: getattr(p, k))
    s = OracleParams()
    for name, _ in OracleParams._fields_:
        key = name
        try:
            v = get(key)
        except (KeyError, AttributeError):
            if name == "lambda_":
                v = get("lam")
            elif name in ("n_threads", "wetting", "fix_particles", "interp_u", "kt", "noise_seed"):
                v = 0
            else:
                raise
        setattr(s, name, v)
    return s


def fft3d(data: np.ndarray, sign: int, n_threads: int = 0) -> np.ndarray:
    """Unnormalised 3-D DFT of a complex (nz, ny, nx) array (FFTW convention)."""
    a = np.ascontiguousarray(data, dtype=np.complex128).copy()
    nz, ny, nx = a.shape
    buf = a.view(np.float64)
    rc = lib().oracle_fft3d(nx, ny, nz, _dp(buf), int(sign), int(n_threads))
    assert rc == 0
    return a


def set_fft_order(order: int) -> None:
    """Factor order of the oracle's FFTs made from now on (0: radix 4 then
    2, the default; 1: radix 2 only).  Round-off differs, the DFT not."""
    lib().oracle_set_fft_order(int(order))


def cell_grid(p):
    s = to_struct(p)
    n3 = (C.c_int * 3)()
    d3 = (C.c_double * 3)()
    lib().oracle_cell_grid(C.byref(s), n3, d3)
    return list(n3), list(d3)


def cells_build(p, pos: np.ndarray):
    s = to_struct(p)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = pos.shape[0]
    nc, _ = cell_grid(p)
    ncell = nc[0] * nc[1] * nc[2]
    cs = np.zeros(ncell + 1, np.int32)
    ids = np.zeros(max(n, 1), np.int32)
    mark = C.c_int(-1)
    rc = lib().oracle_cells_build(C.byref(s), n, _dp(pos), _ip(cs), _ip(ids), C.byref(mark))
    return rc, cs, ids[:n], mark.value


def contact_pairs(p, pos: np.ndarray) -> np.ndarray:
    s = to_struct(p)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = pos.shape[0]
    cnt = lib().oracle_contact_pairs(C.byref(s), n, _dp(pos), 0,
                                     C.cast(None, C.POINTER(C.c_int)))
    if cnt < 0:
        raise RuntimeError("particle outside the sorting range")
    out = np.zeros((max(cnt, 1), 2), np.int32)
    cnt2 = lib().oracle_contact_pairs(C.byref(s), n, _dp(pos), cnt, _ip(out))
    assert cnt2 == cnt
    return out[:cnt]


def contact_forces(p, pos, f_pi=None, it_mark=-1, t=1):
    s = to_struct(p)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = pos.shape[0]
    fpp = np.zeros((n, 3))
    tau = C.c_double(0.0)
    fpi = None if f_pi is None else np.ascontiguousarray(f_pi, dtype=np.float64)
    rc = lib().oracle_contact_forces(C.byref(s), n, _dp(pos), _dp(fpi), it_mark, t,
                                     _dp(fpp), C.byref(tau))
    return rc, fpp, tau.value


def interp(p, grid: np.ndarray, pos: np.ndarray) -> np.ndarray:
    s = to_struct(p)
    g = np.ascontiguousarray(grid, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    out = np.zeros(pos.shape[0])
    lib().oracle_interp(C.byref(s), _dp(g), pos.shape[0], _dp(pos), _dp(out))
    return out


def interp2(p, grid: np.ndarray, pos: np.ndarray) -> np.ndarray:
    """field::interpolate2 (field.cpp:251-281) of a real grid to particles."""
    s = to_struct(p)
    g = np.ascontiguousarray(grid, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    out = np.zeros(pos.shape[0])
    lib().oracle_interp2(C.byref(s), _dp(g), pos.shape[0], _dp(pos), _dp(out))
    return out


def brownian(p, t: int, stage: int, n: int, kt=None, noise_seed=None) -> np.ndarray:
    """element::get_f_rand (element.cpp:109-144) of n particles at step t,
    Heun stage 1 or 2: (n, 3)."""
    s = to_struct(p)
    if kt is not None:
        s.kt = kt
    if noise_seed is not None:
        s.noise_seed = noise_seed
    out = np.zeros((n, 3))
    lib().oracle_brownian(C.byref(s), t, stage, n, _dp(out))
    return out


def drag(p, u: np.ndarray, uf: np.ndarray, radius=None) -> np.ndarray:
    """element::get_f_drag (element.cpp:147-151): -6 pi r mu (u - uf)."""
    s = to_struct(p)
    u = np.ascontiguousarray(u, dtype=np.float64)
    uf = np.ascontiguousarray(uf, dtype=np.float64)
    r = None if radius is None else np.ascontiguousarray(radius, dtype=np.float64)
    out = np.zeros_like(u)
    lib().oracle_drag(C.byref(s), u.shape[0], _dp(u), _dp(uf), _dp(r), _dp(out))
    return out


def spread(p, pos: np.ndarray, force: np.ndarray) -> np.ndarray:
    s = to_struct(p)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    force = np.ascontiguousarray(force, dtype=np.float64)
    out = np.zeros((3, p_nz(p), p_ny(p), p_nx(p)))
    lib().oracle_spread(C.byref(s), pos.shape[0], _dp(pos), _dp(force), _dp(out))
    return out


def f_pi(p, c_loc, dcdx, radius=None, cos_theta=None):
    s = to_struct(p)
    c_loc = np.ascontiguousarray(c_loc, dtype=np.float64)
    dcdx = np.ascontiguousarray(dcdx, dtype=np.float64)
    n = c_loc.shape[0]
    out = np.zeros((n, 3))
    r = None if radius is None else np.ascontiguousarray(radius, dtype=np.float64)
    ct = None if cos_theta is None else np.ascontiguousarray(cos_theta, dtype=np.float64)
    lib().oracle_f_pi(C.byref(s), n, _dp(c_loc), _dp(dcdx), _dp(r), _dp(ct), _dp(out))
    return out


def stokes(p, F: np.ndarray) -> np.ndarray:
    """Stokes block alone (main.cpp:175-179) for a real force F (3, nz, ny, nx)."""
    s = to_struct(p)
    F = np.ascontiguousarray(F, dtype=np.float64)
    out = np.zeros_like(F)
    lib().oracle_stokes(C.byref(s), _dp(F), _dp(out))
    return out


def jitter(i, j, a, r_p):
    return lib().oracle_jitter(i, j, a, r_p)


def p_nx(p):
    return p["nx"] if isinstance(p, dict) else p.nx


def p_ny(p):
    return p["ny"] if isinstance(p, dict) else p.ny


def p_nz(p):
    return p["nz"] if isinstance(p, dict) else p.nz


class Sim:
    """Reference time stepping on the CPU (main.cpp:75-222)."""

    def __init__(self, p, phi0, pos=None, radius=None, cos_theta=None, n_threads=0, wetting=0.0,
                 fix_particles=False):
        """wetting, fix_particles: the calibration model of libfipi's
        adjoint (see oracle_params); the defaults are the reference step."""
        self._s = to_struct(p)
        self._s.n_threads = n_threads
        self._s.wetting = float(wetting)
        self._s.fix_particles = 1 if fix_particles else 0
        self.nx, self.ny, self.nz = self._s.nx, self._s.ny, self._s.nz
        self.n = self._s.n_particles if self._s.switch_particle else 0
        phi0 = np.ascontiguousarray(phi0, dtype=np.float64).reshape(-1)
        assert phi0.size == self.nx * self.ny * self.nz
        pos = None if pos is None else np.ascontiguousarray(pos, dtype=np.float64)
        if self.n:
            assert pos is not None and pos.shape == (self.n, 3)
        r = None if radius is None else np.ascontiguousarray(radius, dtype=np.float64)
        ct = None if cos_theta is None else np.ascontiguousarray(cos_theta, dtype=np.float64)
        self._h = lib().oracle_create(C.byref(self._s), _dp(phi0), _dp(pos), _dp(r), _dp(ct))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def step(self, n=1):
        rc = lib().oracle_step(self._h, n)
        if rc != 0:
            raise RuntimeError("particle outside the sorting range")

    @property
    def t(self):
        return lib().oracle_time(self._h)

    def phi(self):
        out = np.zeros((self.nz, self.ny, self.nx))
        lib().oracle_get_phi(self._h, _dp(out))
        return out

    def u(self):
        out = np.zeros((3, self.nz, self.ny, self.nx))
        lib().oracle_get_u(self._h, _dp(out))
        return out

    def grad_c(self):
        out = np.zeros((3, self.nz, self.ny, self.nx))
        lib().oracle_get_grad_c(self._h, _dp(out))
        return out

    def set_u(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        lib().oracle_set_u(self._h, _dp(u))

    def particles(self):
        n = max(self.n, 1)
        arrs = [np.zeros((n, 3)) for _ in range(6)] + [np.zeros(n)]
        lib().oracle_get_particles(self._h, *[_dp(a) for a in arrs])
        keys = ["p", "p_old", "u", "u_old", "f_pp", "f_pi", "c_loc"]
        return {k: a[: self.n] for k, a in zip(keys, arrs)}

    def tau(self):
        return lib().oracle_get_tau(self._h)

    def particle_fluid(self):
        """uf and f_rand of the last Heun stage, (n, 3) each."""
        n = max(self.n, 1)
        uf, fr = np.zeros((n, 3)), np.zeros((n, 3))
        lib().oracle_get_particle_fluid(self._h, _dp(uf), _dp(fr))
        return {"uf": uf[: self.n], "f_rand": fr[: self.n]}
