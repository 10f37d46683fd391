"""fipi_amd: MI355X-native FIPI time step (Cahn-Hilliard + Stokes +
interface-trapped particles) behind the C ABI of include/fipi.h.

This module is the Python mirror of that ABI (ctypes), used by the tests
and the benchmark.  The compute path is libfipi.so (HIP kernels for
gfx950); there is no CPU fallback: every call fails loudly when the library
or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .params import FP32_MIXED, FP64, Params, config_params, reference_params

__all__ = [
    "FP64", "FP32_MIXED", "Params", "config_params", "reference_params",
    "Simulation", "Ensemble", "Adjoint", "FipiError", "lib", "LIB_PATH", "init_phi", "init_rsa", "init_uniform",
    "init_phi_device", "init_rsa_device", "init_uniform_device",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfipi.so")

ERRORS = {
    -1: "FIPI_ERR_ARG", -2: "FIPI_ERR_HIP", -3: "FIPI_ERR_UNSUPPORTED",
    -4: "FIPI_ERR_RANGE", -5: "FIPI_ERR_STATE", -6: "FIPI_ERR_OVERFLOW",
}

# names exported by libfipi.so (declared in include/fipi.h)
ABI_SYMBOLS = [
    "fipi_create", "fipi_destroy", "fipi_last_error", "fipi_stream",
    "fipi_set_state", "fipi_set_velocity", "fipi_get_phi", "fipi_get_velocity",
    "fipi_get_grad_phi", "fipi_get_particles", "fipi_get_time", "fipi_diagnostics",
    "fipi_step", "fipi_step_async", "fipi_sync", "fipi_profile_steps",
    "fipi_fft_r2c", "fipi_fft_c2r", "fipi_cells_build", "fipi_cells_build_noalloc", "fipi_contact_pairs",
    "fipi_contact_forces", "fipi_interp", "fipi_spread", "fipi_interp2", "fipi_brownian_forces",
    "fipi_get_particle_fluid",
    "fipi_buffer_create", "fipi_buffer_destroy", "fipi_buffer_upload", "fipi_buffer_download", "fipi_field_fft",
    "fipi_field_op", "fipi_capillary_forces", "fipi_advect",
    "fipi_snapshot_size", "fipi_snapshot_save", "fipi_snapshot_load", "fipi_tune", "fipi_get_tuning",
    "fipi_set_tuning",
    "fipi_ensemble_create", "fipi_ensemble_destroy", "fipi_ensemble_last_error", "fipi_ensemble_set_box",
    "fipi_ensemble_tune", "fipi_ensemble_step", "fipi_ensemble_step_async", "fipi_ensemble_sync",
    "fipi_ensemble_stream", "fipi_ensemble_get_phi", "fipi_ensemble_get_particles",
    "fipi_ensemble_diagnostics", "fipi_ensemble_get_tuning",
    "fipi_ensemble_set_tuning",
    "fipi_adjoint_create", "fipi_adjoint_destroy", "fipi_adjoint_last_error", "fipi_adjoint_set_target",
    "fipi_adjoint_forward", "fipi_adjoint_gradient",
    "fipi_init_phi", "fipi_init_rsa", "fipi_init_uniform",
    "fipi_init_phi_device", "fipi_init_uniform_device", "fipi_init_rsa_device", "fipi_init_last_error",
    "fipi_set_state_device",
]


class FipiError(RuntimeError):
    pass


class CKernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 40), ("launches", C.c_int32), ("total_ms", C.c_double),
                ("bytes", C.c_double)]


class CParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("n_particles", C.c_int32),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("pi", C.c_double), ("dt", C.c_double), ("ramp_t", C.c_double),
        ("epsilon", C.c_double), ("lambda_", C.c_double), ("mobility", C.c_double),
        ("mu", C.c_double), ("sigma", C.c_double), ("a_pi", C.c_double),
        ("k_n", C.c_double), ("r_p", C.c_double), ("r_eq", C.c_double),
        ("r_c", C.c_double), ("margin", C.c_double),
        ("rho1", C.c_double), ("rho2", C.c_double), ("gravity", C.c_double),
        ("tanh2", C.c_double),
        ("switch_particle", C.c_int32), ("switch_ch", C.c_int32), ("switch_ns", C.c_int32),
        ("precision", C.c_int32),
        ("interp_u", C.c_int32), ("kt", C.c_double), ("noise_seed", C.c_uint64),
    ]


class CBoxDesc(C.Structure):
    _fields_ = [("n", C.c_int32), ("n_particles", C.c_int32)]


class CDiag(C.Structure):
    _fields_ = [
        ("t", C.c_int32), ("it_mark", C.c_int32), ("mean_phi", C.c_double),
        ("free_energy", C.c_double), ("interface_area", C.c_double), ("tau", C.c_double),
    ]


def _to_cparams(p: Params) -> CParams:
    s = CParams()
    for name, _ in CParams._fields_:
        setattr(s, name, getattr(p, "lam" if name == "lambda_" else name))
    return s


_lib = None


class CTuning(C.Structure):
    """fipi_tuning_t: tile widths in bytes and the y-pass FFT plan (values
    per thread; 0 = default when set)."""
    _fields_ = [("y_tile_bytes", C.c_int32), ("hub_tile_bytes", C.c_int32),
                ("stokes_tile_bytes", C.c_int32), ("y_values_per_thread", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}

    @classmethod
    def from_dict(cls, d: dict) -> "CTuning":
        return cls(*(int(d.get(k, 0)) for k, _ in cls._fields_))


def lib():
    """Load libfipi.so (built in-tree by fipi_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FipiError(f"{LIB_PATH} is missing: run `python -m fipi_amd.build` first")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
    i64 = C.c_int64
    L.fipi_create.argtypes = [C.POINTER(CParams), C.POINTER(vp)]
    L.fipi_destroy.argtypes = [vp]
    L.fipi_last_error.argtypes = [vp]
    L.fipi_last_error.restype = C.c_char_p
    L.fipi_stream.argtypes = [vp]
    L.fipi_stream.restype = vp
    L.fipi_set_state.argtypes = [vp, dp, dp, dp, dp]
    L.fipi_set_velocity.argtypes = [vp, dp]
    for nm in ("fipi_get_phi", "fipi_get_velocity", "fipi_get_grad_phi"):
        getattr(L, nm).argtypes = [vp, dp]
    L.fipi_get_particles.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.fipi_get_time.argtypes = [vp, ip]
    L.fipi_diagnostics.argtypes = [vp, C.POINTER(CDiag)]
    L.fipi_step.argtypes = [vp, C.c_int32]
    L.fipi_step_async.argtypes = [vp, C.c_int32]
    L.fipi_sync.argtypes = [vp]
    L.fipi_profile_steps.argtypes = [vp, C.c_int32, C.POINTER(CKernelStat), C.c_int32, ip]
    L.fipi_fft_r2c.argtypes = [vp, vp, vp]
    L.fipi_fft_c2r.argtypes = [vp, vp, vp]
    L.fipi_cells_build.argtypes = [vp, vp, C.c_int32, vp, vp, ip]
    L.fipi_cells_build_noalloc.argtypes = [C.POINTER(CParams), vp, C.c_int32, vp, vp, vp, ip, vp]
    L.fipi_buffer_create.argtypes = [vp, i64, C.POINTER(vp)]
    L.fipi_buffer_destroy.argtypes = [vp, vp]
    L.fipi_buffer_upload.argtypes = [vp, vp, vp, i64]
    L.fipi_buffer_download.argtypes = [vp, vp, vp, i64]
    L.fipi_field_fft.argtypes = [vp, vp, C.c_int32, C.c_int32]
    L.fipi_field_op.argtypes = [vp, C.c_int32, vp, C.c_int32, vp, C.c_int32, C.c_int32, vp, C.c_int32, C.c_int32,
                                C.c_double]
    L.fipi_capillary_forces.argtypes = [vp, vp, vp, vp, vp, C.c_int32, vp]
    L.fipi_advect.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp, vp]
    L.fipi_contact_pairs.argtypes = [vp, vp, C.c_int32, vp, i64, C.POINTER(i64)]
    L.fipi_contact_forces.argtypes = [vp, vp, C.c_int32, C.c_int32, vp]
    L.fipi_interp.argtypes = [vp, vp, vp, C.c_int32, vp]
    L.fipi_spread.argtypes = [vp, vp, vp, C.c_int32, vp]
    L.fipi_interp2.argtypes = [vp, vp, vp, C.c_int32, vp]
    L.fipi_brownian_forces.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, vp]
    L.fipi_get_particle_fluid.argtypes = [vp, vp, vp, vp]
    L.fipi_init_phi.argtypes = [i64, C.c_double, C.c_double, C.c_uint64, dp]
    L.fipi_init_rsa.argtypes = [i64, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_uint64, i64, dp]
    L.fipi_init_rsa.restype = i64
    L.fipi_init_uniform.argtypes = [i64, C.c_double, C.c_double, C.c_uint64, dp]
    L.fipi_init_phi_device.argtypes = [i64, C.c_double, C.c_double, C.c_uint64, vp, vp]
    L.fipi_init_uniform_device.argtypes = [i64, C.c_double, C.c_double, C.c_uint64, vp, vp]
    L.fipi_init_rsa_device.argtypes = [i64, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_uint64, i64, vp, vp]
    L.fipi_init_rsa_device.restype = i64
    L.fipi_init_last_error.argtypes = []
    L.fipi_init_last_error.restype = C.c_char_p
    L.fipi_set_state_device.argtypes = [vp, vp, vp, vp, vp]
    L.fipi_snapshot_size.argtypes = [vp]
    L.fipi_snapshot_size.restype = i64
    L.fipi_snapshot_save.argtypes = [vp, vp, i64]
    L.fipi_snapshot_load.argtypes = [vp, vp, i64]
    L.fipi_tune.argtypes = [vp, C.c_int32]
    L.fipi_get_tuning.argtypes = [vp, C.POINTER(CTuning)]
    L.fipi_set_tuning.argtypes = [vp, C.POINTER(CTuning)]
    L.fipi_ensemble_create.argtypes = [C.POINTER(CParams), C.c_int32, C.POINTER(CBoxDesc), C.POINTER(vp)]
    L.fipi_ensemble_destroy.argtypes = [vp]
    L.fipi_ensemble_last_error.argtypes = [vp]
    L.fipi_ensemble_last_error.restype = C.c_char_p
    L.fipi_ensemble_set_box.argtypes = [vp, C.c_int32, dp, dp, dp, dp]
    L.fipi_ensemble_tune.argtypes = [vp, C.c_int32]
    L.fipi_ensemble_step.argtypes = [vp, C.c_int32]
    L.fipi_ensemble_step_async.argtypes = [vp, C.c_int32]
    L.fipi_ensemble_sync.argtypes = [vp]
    L.fipi_ensemble_stream.argtypes = [vp]
    L.fipi_ensemble_stream.restype = vp
    L.fipi_ensemble_get_phi.argtypes = [vp, C.c_int32, dp]
    L.fipi_ensemble_get_particles.argtypes = [vp, C.c_int32, dp]
    L.fipi_ensemble_diagnostics.argtypes = [vp, C.c_int32, C.POINTER(CDiag)]
    L.fipi_ensemble_get_tuning.argtypes = [vp, C.c_int32, C.POINTER(CTuning)]
    L.fipi_ensemble_set_tuning.argtypes = [vp, C.c_int32, C.POINTER(CTuning)]
    L.fipi_adjoint_create.argtypes = [C.POINTER(CParams), dp, dp, C.c_double, C.c_int32, C.c_int32, C.POINTER(vp)]
    L.fipi_adjoint_destroy.argtypes = [vp]
    L.fipi_adjoint_last_error.argtypes = [vp]
    L.fipi_adjoint_last_error.restype = C.c_char_p
    L.fipi_adjoint_set_target.argtypes = [vp, dp]
    L.fipi_adjoint_forward.argtypes = [vp, C.c_double, C.c_double, dp, dp, dp]
    L.fipi_adjoint_gradient.argtypes = [vp, C.c_double, C.c_double, dp, dp, dp, dp, dp]
    _lib = L
    return L


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc, h=None):
    if rc != 0:
        msg = lib().fipi_last_error(h)
        raise FipiError(f"{ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")


# ---- seeded synthetic inputs (host only) ---------------------------------
def init_phi(shape, mean: float, amp: float, seed: int) -> np.ndarray:
    out = np.empty(int(np.prod(shape)), np.float64)
    _check(lib().fipi_init_phi(out.size, mean, amp, seed, _dp(out)))
    return out.reshape(shape)


def init_rsa(n: int, r: float, gap: float, box, seed: int, max_attempts: int = 100000) -> np.ndarray:
    out = np.zeros((max(n, 1), 3), np.float64)
    placed = lib().fipi_init_rsa(n, r, gap, box[0], box[1], box[2], seed, max_attempts, _dp(out))
    if placed < 0:
        _check(int(placed))
    return out[:placed].copy()


def init_uniform(n: int, lo: float, hi: float, seed: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    _check(lib().fipi_init_uniform(n, lo, hi, seed, _dp(out)))
    return out


def _dev_init_check(rc):
    if rc < 0:
        raise FipiError(f"{ERRORS.get(int(rc), rc)}: {lib().fipi_init_last_error().decode()}")
    return rc


def init_phi_device(shape, mean: float, amp: float, seed: int):
    """init_phi's values generated on the GPU (a float64 CUDA tensor)."""
    import torch
    out = torch.empty(int(np.prod(shape)), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _dev_init_check(lib().fipi_init_phi_device(out.numel(), mean, amp, seed, C.c_void_p(out.data_ptr()), None))
    return out.reshape(tuple(shape))


def init_uniform_device(n: int, lo: float, hi: float, seed: int):
    import torch
    out = torch.empty(max(n, 1), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _dev_init_check(lib().fipi_init_uniform_device(n, lo, hi, seed, C.c_void_p(out.data_ptr()), None))
    return out[:n]


def init_rsa_device(n: int, r: float, gap: float, box, seed: int, max_attempts: int = 100000):
    """init_rsa's positions generated on the GPU: an (placed, 3) float64
    CUDA tensor."""
    import torch
    out = torch.zeros((max(n, 1), 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    placed = _dev_init_check(lib().fipi_init_rsa_device(n, r, gap, box[0], box[1], box[2], seed, max_attempts,
                                                        C.c_void_p(out.data_ptr()), None))
    return out[:placed]


class Simulation:
    """One periodic box on the current GPU (a fipi_ctx)."""

    def __init__(self, params: Params):
        self.params = params
        self._h = C.c_void_p()
        self._cp = _to_cparams(params)
        _check(lib().fipi_create(C.byref(self._cp), C.byref(self._h)))
        self.n = params.n_particles if params.switch_particle else 0

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fipi_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        _check(rc, self._h)

    @property
    def handle(self):
        return self._h

    def set_state(self, phi, pos=None, radius=None, cos_theta=None):
        p = self.params
        phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
        assert phi.size == p.G
        if self.n:
            pos = np.ascontiguousarray(pos, dtype=np.float64)
            assert pos.shape == (self.n, 3)
        r = None if radius is None else np.ascontiguousarray(radius, dtype=np.float64)
        ct = None if cos_theta is None else np.ascontiguousarray(cos_theta, dtype=np.float64)
        self._c(lib().fipi_set_state(self._h, _dp(phi), _dp(pos) if self.n else None, _dp(r), _dp(ct)))

    def set_state_device(self, phi, pos=None, radius=None, cos_theta=None):
        """set_state from float64 CUDA tensors (no host copy of the state)."""
        import torch
        p = self.params
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        phi = phi.contiguous()
        assert phi.dtype == torch.float64 and phi.numel() == p.G
        if self.n:
            pos = pos.contiguous()
            assert pos.dtype == torch.float64 and tuple(pos.shape) == (self.n, 3)
        r = None if radius is None else radius.contiguous()
        ct = None if cos_theta is None else cos_theta.contiguous()
        torch.cuda.synchronize()
        self._c(lib().fipi_set_state_device(self._h, ptr(phi), ptr(pos) if self.n else None, ptr(r), ptr(ct)))

    def set_velocity(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.size == 3 * self.params.G
        self._c(lib().fipi_set_velocity(self._h, _dp(u)))

    def step(self, n: int = 1):
        self._c(lib().fipi_step(self._h, n))

    def step_async(self, n: int = 1):
        self._c(lib().fipi_step_async(self._h, n))

    def sync(self):
        self._c(lib().fipi_sync(self._h))

    def profile_steps(self, n: int = 1) -> list:
        """Run n steps with per-launch HIP events; returns one dict per launch
        site: name, launches, total_ms, bytes (algorithmic)."""
        arr = (CKernelStat * 64)()
        cnt = C.c_int32()
        self._c(lib().fipi_profile_steps(self._h, n, arr, 64, C.byref(cnt)))
        return [{"name": arr[i].name.decode(), "launches": arr[i].launches,
                 "total_ms": arr[i].total_ms, "bytes": arr[i].bytes} for i in range(min(cnt.value, 64))]

    @property
    def stream_ptr(self) -> int:
        return lib().fipi_stream(self._h) or 0

    def phi(self) -> np.ndarray:
        p = self.params
        out = np.empty((p.nz, p.ny, p.nx))
        self._c(lib().fipi_get_phi(self._h, _dp(out)))
        return out

    def velocity(self) -> np.ndarray:
        p = self.params
        out = np.empty((3, p.nz, p.ny, p.nx))
        self._c(lib().fipi_get_velocity(self._h, _dp(out)))
        return out

    def grad_phi(self) -> np.ndarray:
        p = self.params
        out = np.empty((3, p.nz, p.ny, p.nx))
        self._c(lib().fipi_get_grad_phi(self._h, _dp(out)))
        return out

    def particles(self) -> dict:
        n = max(self.n, 1)
        arrs = [np.zeros((n, 3)) for _ in range(6)]
        self._c(lib().fipi_get_particles(self._h, *[_dp(a) for a in arrs]))
        keys = ["p", "p_old", "u", "u_old", "f_pp", "f_pi"]
        return {k: a[: self.n] for k, a in zip(keys, arrs)}

    def particle_fluid(self) -> dict:
        """uf, f_rand and f_drag = -6 pi r mu (u - uf) after the last Heun
        stage, (n, 3) each (zero unless interp_u / kt are set)."""
        n = max(self.n, 1)
        arrs = [np.zeros((n, 3)) for _ in range(3)]
        self._c(lib().fipi_get_particle_fluid(self._h, *[_dp(a) for a in arrs]))
        return {k: a[: self.n] for k, a in zip(["uf", "f_rand", "f_drag"], arrs)}

    @property
    def t(self) -> int:
        v = C.c_int32()
        self._c(lib().fipi_get_time(self._h, C.byref(v)))
        return v.value

    def diagnostics(self) -> dict:
        d = CDiag()
        self._c(lib().fipi_diagnostics(self._h, C.byref(d)))
        return {k: getattr(d, k) for k, _ in CDiag._fields_}

    def snapshot(self) -> bytes:
        """The whole state (fields, SBDF2 history, particles, counters)."""
        n = lib().fipi_snapshot_size(self._h)
        if n < 0:
            self._c(int(n))
        buf = C.create_string_buffer(int(n))
        self._c(lib().fipi_snapshot_save(self._h, buf, n))
        return buf.raw

    def restore(self, snap: bytes):
        """Continue from a snapshot of a context with the same parameters."""
        buf = C.create_string_buffer(snap, len(snap))
        self._c(lib().fipi_snapshot_load(self._h, buf, len(snap)))

    def tune(self, steps: int = 3) -> dict:
        """Pick the fastest tile widths and y-pass FFT plan over `steps`
        steps each; the state is left as it was.  Returns the choice."""
        self._c(lib().fipi_tune(self._h, steps))
        return self.tuning()

    def tuning(self) -> dict:
        t = CTuning()
        self._c(lib().fipi_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, tuning: dict):
        self._c(lib().fipi_set_tuning(self._h, C.byref(CTuning.from_dict(tuning))))

    # ---- unit entry points on device memory (torch tensors) --------------
    # torch only provides device memory here; the work runs in libfipi.
    def _sync_torch(self):
        import torch
        torch.cuda.synchronize()

    def fft_r2c(self, real):
        import torch
        p = self.params
        dt = torch.float64 if p.precision == FP64 else torch.float32
        assert real.dtype == dt and real.is_contiguous() and real.numel() == p.G
        out = torch.empty((p.nz, p.ny, p.nx // 2 + 1, 2), dtype=dt, device=real.device)
        self._sync_torch()
        self._c(lib().fipi_fft_r2c(self._h, C.c_void_p(real.data_ptr()), C.c_void_p(out.data_ptr())))
        return out

    def fft_c2r(self, half):
        import torch
        p = self.params
        half = half.clone()
        out = torch.empty((p.nz, p.ny, p.nx), dtype=half.dtype, device=half.device)
        self._sync_torch()
        self._c(lib().fipi_fft_c2r(self._h, C.c_void_p(half.data_ptr()), C.c_void_p(out.data_ptr())))
        return out

    def cells_build(self, pos):
        import torch
        pos = torch.as_tensor(np.ascontiguousarray(pos, dtype=np.float64)).cuda()
        n = pos.shape[0]
        nc, _ = self.params.cell_grid()
        ncell = nc[0] * nc[1] * nc[2]
        start = torch.zeros(ncell + 1, dtype=torch.int32, device=pos.device)
        ids = torch.zeros(max(n, 1), dtype=torch.int32, device=pos.device)
        ncell_out = C.c_int32()
        self._sync_torch()
        self._c(lib().fipi_cells_build(self._h, C.c_void_p(pos.data_ptr()), n,
                                       C.c_void_p(start.data_ptr()), C.c_void_p(ids.data_ptr()),
                                       C.byref(ncell_out)))
        assert ncell_out.value == ncell
        return start.cpu().numpy(), ids[:n].cpu().numpy()

    def contact_pairs(self, pos):
        import torch
        pos = torch.as_tensor(np.ascontiguousarray(pos, dtype=np.float64)).cuda()
        n = pos.shape[0]
        cnt = C.c_int64()
        self._sync_torch()
        self._c(lib().fipi_contact_pairs(self._h, C.c_void_p(pos.data_ptr()), n, None, 0, C.byref(cnt)))
        m = cnt.value
        pairs = torch.zeros((max(m, 1), 2), dtype=torch.int32, device=pos.device)
        self._c(lib().fipi_contact_pairs(self._h, C.c_void_p(pos.data_ptr()), n,
                                         C.c_void_p(pairs.data_ptr()), m, C.byref(cnt)))
        assert cnt.value == m
        return pairs[:m].cpu().numpy()

    def contact_forces(self, pos, t: int):
        import torch
        pos = torch.as_tensor(np.ascontiguousarray(pos, dtype=np.float64)).cuda()
        n = pos.shape[0]
        out = torch.zeros((max(n, 1), 3), dtype=torch.float64, device=pos.device)
        self._sync_torch()
        self._c(lib().fipi_contact_forces(self._h, C.c_void_p(pos.data_ptr()), n, t,
                                          C.c_void_p(out.data_ptr())))
        return out[:n].cpu().numpy()

    def interp(self, grid, pos):
        import torch
        p = self.params
        dt = torch.float64 if p.precision == FP64 else torch.float32
        g = torch.as_tensor(np.ascontiguousarray(grid)).to(dt).cuda().contiguous()
        pos = torch.as_tensor(np.ascontiguousarray(pos, dtype=np.float64)).cuda()
        n = pos.shape[0]
        out = torch.zeros(max(n, 1), dtype=torch.float64, device=pos.device)
        self._sync_torch()
        self._c(lib().fipi_interp(self._h, C.c_void_p(g.data_ptr()), C.c_void_p(pos.data_ptr()), n,
                                  C.c_void_p(out.data_ptr())))
        return out[:n].cpu().numpy()

    def interp2(self, grid, pos):
        import torch
        p = self.params
        dt = torch.float64 if p.precision == FP64 else torch.float32
        g = torch.as_tensor(np.ascontiguousarray(grid)).to(dt).cuda().contiguous()
        pos = torch.as_tensor(np.ascontiguousarray(pos, dtype=np.float64)).cuda()
        n = pos.shape[0]
        out = torch.zeros(max(n, 1), dtype=torch.float64, device=pos.device)
        self._sync_torch()
        self._c(lib().fipi_interp2(self._h, C.c_void_p(g.data_ptr()), C.c_void_p(pos.data_ptr()), n,
                                   C.c_void_p(out.data_ptr())))
        return out[:n].cpu().numpy()

    def brownian(self, t, stage, n):
        import torch
        out = torch.zeros((max(n, 1), 3), dtype=torch.float64, device="cuda")
        self._sync_torch()
        self._c(lib().fipi_brownian_forces(self._h, t, stage, n, C.c_void_p(out.data_ptr())))
        return out[:n].cpu().numpy()

    def spread(self, pos, force):
        import torch
        p = self.params
        dt = torch.float64 if p.precision == FP64 else torch.float32
        pos = torch.as_tensor(np.ascontiguousarray(pos, dtype=np.float64)).cuda()
        f = torch.as_tensor(np.ascontiguousarray(force, dtype=np.float64)).cuda()
        n = pos.shape[0]
        grid = torch.empty((3, p.nz, p.ny, p.nx), dtype=dt, device=pos.device)
        self._sync_torch()
        self._c(lib().fipi_spread(self._h, C.c_void_p(pos.data_ptr()), C.c_void_p(f.data_ptr()), n,
                                  C.c_void_p(grid.data_ptr())))
        return grid.cpu().numpy()


class Ensemble:
    """Many independent periodic cubic boxes advanced together (a
    fipi_ensemble): common physics `params` (its lx/nx is the grid spacing
    of every box), per-box edge lengths and particle counts."""

    def __init__(self, params: Params, edges, n_particles):
        self.params = params
        self.edges = [int(n) for n in edges]
        self.counts = [int(k) for k in n_particles]
        assert len(self.edges) == len(self.counts)
        desc = (CBoxDesc * len(self.edges))(*[CBoxDesc(n, k) for n, k in zip(self.edges, self.counts)])
        self._h = C.c_void_p()
        self._cp = _to_cparams(params)
        rc = lib().fipi_ensemble_create(C.byref(self._cp), len(self.edges), desc, C.byref(self._h))
        if rc != 0:
            raise FipiError(f"{ERRORS.get(rc, rc)}: {lib().fipi_ensemble_last_error(None).decode()}")

    def _c(self, rc):
        if rc != 0:
            msg = lib().fipi_ensemble_last_error(self._h)
            raise FipiError(f"{ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fipi_ensemble_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_box(self, b, phi, pos=None, radius=None, cos_theta=None):
        n, k = self.edges[b], self.counts[b]
        phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
        assert phi.size == n ** 3
        if k:
            pos = np.ascontiguousarray(pos, dtype=np.float64)
            assert pos.shape == (k, 3)
        r = None if radius is None else np.ascontiguousarray(radius, dtype=np.float64)
        ct = None if cos_theta is None else np.ascontiguousarray(cos_theta, dtype=np.float64)
        self._c(lib().fipi_ensemble_set_box(self._h, b, _dp(phi), _dp(pos) if k else None, _dp(r), _dp(ct)))

    def tune(self, steps=3):
        self._c(lib().fipi_ensemble_tune(self._h, steps))

    def tuning(self, n) -> dict:
        """Choices of the slab of edge n (all zero when there is none)."""
        t = CTuning()
        self._c(lib().fipi_ensemble_get_tuning(self._h, n, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, n, tuning: dict):
        self._c(lib().fipi_ensemble_set_tuning(self._h, n, C.byref(CTuning.from_dict(tuning))))

    def step(self, nsteps=1):
        self._c(lib().fipi_ensemble_step(self._h, nsteps))

    def step_async(self, nsteps=1):
        self._c(lib().fipi_ensemble_step_async(self._h, nsteps))

    def sync(self):
        self._c(lib().fipi_ensemble_sync(self._h))

    @property
    def stream_ptr(self):
        return lib().fipi_ensemble_stream(self._h)

    def phi(self, b):
        n = self.edges[b]
        out = np.empty(n ** 3, np.float64)
        self._c(lib().fipi_ensemble_get_phi(self._h, b, _dp(out)))
        return out.reshape(n, n, n)

    def particles(self, b):
        out = np.zeros((max(self.counts[b], 1), 3), np.float64)
        self._c(lib().fipi_ensemble_get_particles(self._h, b, _dp(out)))
        return out[:self.counts[b]]

    def diagnostics(self, b):
        d = CDiag()
        self._c(lib().fipi_ensemble_diagnostics(self._h, b, C.byref(d)))
        return {k: getattr(d, k) for k, _ in CDiag._fields_}


class Adjoint:
    """Calibration problem with a hand-written adjoint (a fipi_adjoint):
    particles held at `pos`, wetting strength `wetting` (a_w), loss
    sum (phi_T - phi_target)^2 after `steps` steps from phi0, forward state
    checkpointed every `checkpoint_every` steps."""

    def __init__(self, params: Params, phi0, pos, wetting: float, steps: int, checkpoint_every: int = 20):
        self.params = params
        self.n = params.n_particles if params.switch_particle else 0
        phi0 = np.ascontiguousarray(phi0, dtype=np.float64).reshape(-1)
        assert phi0.size == params.G
        if self.n:
            pos = np.ascontiguousarray(pos, dtype=np.float64)
            assert pos.shape == (self.n, 3)
        self._h = C.c_void_p()
        self._cp = _to_cparams(params)
        rc = lib().fipi_adjoint_create(C.byref(self._cp), _dp(phi0), _dp(pos) if self.n else None, wetting, steps,
                                       checkpoint_every, C.byref(self._h))
        if rc != 0:
            raise FipiError(f"{ERRORS.get(rc, rc)}: {lib().fipi_adjoint_last_error(None).decode()}")

    def _c(self, rc):
        if rc != 0:
            msg = lib().fipi_adjoint_last_error(self._h)
            raise FipiError(f"{ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().fipi_adjoint_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _theta(self, theta):
        if not self.n:
            return None
        th = np.ascontiguousarray(theta, dtype=np.float64)
        assert th.shape == (self.n,)
        return th

    def set_target(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
        assert phi.size == self.params.G
        self._c(lib().fipi_adjoint_set_target(self._h, _dp(phi)))

    def forward(self, mobility, lam, theta, want_loss=True):
        """(loss or None, final phi) at the given parameters."""
        p = self.params
        out = np.empty(p.G, np.float64)
        loss = np.zeros(1)
        self._c(lib().fipi_adjoint_forward(self._h, mobility, lam, _dp(self._theta(theta)),
                                           _dp(loss) if want_loss else None, _dp(out)))
        return (float(loss[0]) if want_loss else None), out.reshape(p.nz, p.ny, p.nx)

    def gradient(self, mobility, lam, theta):
        """(loss, dloss/dmobility, dloss/dlambda, dloss/dtheta)."""
        loss, gm, gl = np.zeros(1), np.zeros(1), np.zeros(1)
        gth = np.zeros(max(self.n, 1))
        self._c(lib().fipi_adjoint_gradient(self._h, mobility, lam, _dp(self._theta(theta)), _dp(loss), _dp(gm),
                                            _dp(gl), _dp(gth) if self.n else None))
        return float(loss[0]), float(gm[0]), float(gl[0]), gth[:self.n].copy()
