"""CPU checks of the boundary: libfipi.so loads without a GPU, exports every
function include/fipi.h declares, the struct layouts match, and host-side
argument checking returns status codes instead of exiting."""
import ctypes as C
import os
import re

import fipi_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "fipi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fipi_[a-z0-9_]+)\s*\(", src)))


def test_header_and_python_mirror_agree():
    assert _declared() == sorted(fipi_amd.ABI_SYMBOLS)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(fipi_amd.LIB_PATH)
    for name in _declared():
        assert hasattr(L, name), name


def test_struct_sizes():
    # 4 int32 + 21 doubles + 4 int32, interp_u and its padding, kt, noise_seed
    # (include/fipi.h; fipi.hip asserts the same size)
    assert C.sizeof(fipi_amd.CParams) == 4 * 4 + 21 * 8 + 4 * 4 + 8 + 8 + 8
    assert C.sizeof(fipi_amd.CDiag) == 8 + 4 * 8
    assert C.sizeof(fipi_amd.CKernelStat) == 40 + 4 + 4 + 8 + 8


def test_param_checks_return_codes():
    L = fipi_amd.lib()
    h = C.c_void_p()
    bad = fipi_amd._to_cparams(fipi_amd.config_params(40, 0, 2.0, 10))
    assert L.fipi_create(C.byref(bad), C.byref(h)) == -3  # FIPI_ERR_UNSUPPORTED
    assert b"grid sizes" in L.fipi_last_error(None)
    p = fipi_amd.config_params(32, 0, 2.0, 10).replace(switch_ch=0)
    assert L.fipi_create(C.byref(fipi_amd._to_cparams(p)), C.byref(h)) == -3
    p = fipi_amd.config_params(32, 10, 2.0, 10).replace(r_c=12.0)  # 2 cells per axis
    assert L.fipi_create(C.byref(fipi_amd._to_cparams(p)), C.byref(h)) == -1
    assert L.fipi_create(None, C.byref(h)) == -1
    p = fipi_amd.config_params(32, 10, 2.0, 10).replace(interp_u=3)
    assert L.fipi_create(C.byref(fipi_amd._to_cparams(p)), C.byref(h)) == -1
    p = fipi_amd.config_params(32, 10, 2.0, 10).replace(kt=-1.0)
    assert L.fipi_create(C.byref(fipi_amd._to_cparams(p)), C.byref(h)) == -1
    assert L.fipi_step(None, 1) == -1
    assert L.fipi_destroy(None) == 0


def test_reference_params_match_head_hpp():
    """head.hpp:27-74: 64^3, L = 2*3.1415926, dt = 1e-3, eps = 3dx/sqrt2/2,
    lambda = sigma eps 3/(2 sqrt2), M = (0.2 eps)^2/mu, r_p = 0.03*0.4 L,
    r_eq = r_c = 5 r_p, k_n = sigma/2."""
    p = fipi_amd.reference_params()
    dx = 2 * 3.1415926 / 64
    assert p.nx == p.ny == p.nz == 64 and p.dt == 1e-3
    assert abs(p.epsilon - 3 * dx / 2 ** 0.5 / 2) < 1e-17
    assert abs(p.lam - 0.07 * p.epsilon * 3 / 2 / 2 ** 0.5) < 1e-17
    assert abs(p.mobility - (0.2 * p.epsilon) ** 2 / 1e-3) < 1e-15
    assert abs(p.r_p - 0.03 * 0.4 * 2 * 3.1415926) < 1e-16
    assert p.r_eq == p.r_c == 5 * p.r_p and p.k_n == 0.5 * 0.07
    n, d = p.cell_grid()
    assert n == [16, 16, 16]


def test_tuning_struct_round_trip():
    # fipi_tuning_t: four int32 (tile bytes of the y passes, CH hub, Stokes
    # hub, and the y-pass values per thread)
    assert C.sizeof(fipi_amd.CTuning) == 4 * 4
    d = {"y_tile_bytes": 64, "hub_tile_bytes": 128, "stokes_tile_bytes": 64, "y_values_per_thread": 16}
    assert fipi_amd.CTuning.from_dict(d).as_dict() == d
    assert fipi_amd.CTuning.from_dict({}).as_dict() == {k: 0 for k in d}


def test_tuning_calls_reject_null_context():
    L = fipi_amd.lib()
    t = fipi_amd.CTuning()
    assert L.fipi_get_tuning(None, C.byref(t)) == -1
    assert L.fipi_set_tuning(None, C.byref(t)) == -1


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
