"""fipi_cells_build_noalloc (the cell list without any allocation) against
fipi_cells_build on the same inputs, through the C ABI: bit-exact cell
starts and ids on two calls in a row, the second with a smaller input, and
the oracle's cell list for the first."""
import ctypes as C

import numpy as np
import pytest

import fipi_amd
import oracle
from fipi_amd import config_params

pytestmark = pytest.mark.gpu


def _noalloc(p, pos):
    import torch
    L = fipi_amd.lib()
    n = pos.shape[0]
    nc, _ = p.cell_grid()
    ncell = nc[0] * nc[1] * nc[2]
    d_pos = torch.as_tensor(np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)).cuda()
    # output arrays poisoned so that every entry must be written
    start = torch.full((ncell + 1,), -7, dtype=torch.int32, device="cuda")
    ids = torch.full((max(n, 1),), -7, dtype=torch.int32, device="cuda")
    err = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    nco = C.c_int32(-1)
    cp = fipi_amd._to_cparams(p)
    torch.cuda.synchronize()
    rc = L.fipi_cells_build_noalloc(C.byref(cp), C.c_void_p(d_pos.data_ptr()), n,
                                    C.c_void_p(start.data_ptr()), C.c_void_p(ids.data_ptr()),
                                    C.c_void_p(err.data_ptr()), C.byref(nco), None)
    assert rc == 0, L.fipi_last_error(None)
    torch.cuda.synchronize()
    assert nco.value == ncell
    return start.cpu().numpy(), ids[:n].cpu().numpy(), int(err.item())


def _with_ctx(p, pos):
    sim = fipi_amd.Simulation(p)
    try:
        return sim.cells_build(pos)
    finally:
        sim.close()


def test_noalloc_matches_cells_build_twice(gpu):
    print("noalloc cells: 128^3 + 12500, then 32^3 + 150", flush=True)
    # first call: the headline's sphere density at 128^3
    p1 = config_params(128, 12500, 2.0, 100)
    pos1 = fipi_amd.init_rsa(12500, 2.0, 0.5, (p1.lx, p1.ly, p1.lz), 2)
    s1, i1, e1 = _noalloc(p1, pos1)
    r1 = _with_ctx(p1, pos1)
    assert e1 == 0
    assert np.array_equal(s1, r1[0]) and np.array_equal(i1, r1[1])
    rc, os_, oids, _ = oracle.cells_build(p1, pos1)
    assert rc == 0 and np.array_equal(s1, os_) and np.array_equal(i1, oids)
    # second call, smaller: fewer particles and a smaller contact grid; the
    # shared scratch keeps the first call's contents and must not leak in
    p2 = config_params(32, 150, 2.0, 100)
    pos2 = fipi_amd.init_rsa(150, 2.0, 0.5, (p2.lx, p2.ly, p2.lz), 5)
    s2, i2, e2 = _noalloc(p2, pos2)
    r2 = _with_ctx(p2, pos2)
    assert e2 == 0
    assert np.array_equal(s2, r2[0]) and np.array_equal(i2, r2[1])


def test_noalloc_limits_and_errors(gpu):
    print("noalloc cells: limits", flush=True)
    L = fipi_amd.lib()
    # beyond the stated maximum: refused on the host before any GPU work
    p = config_params(256, 300000, 2.0, 100)
    cp = fipi_amd._to_cparams(p)
    dummy = C.c_void_p(16)
    nco = C.c_int32()
    rc = L.fipi_cells_build_noalloc(C.byref(cp), dummy, 300000, dummy, dummy, dummy, C.byref(nco), None)
    assert rc == -1 and b"FIPI_NOALLOC_MAX_PARTICLES" in L.fipi_last_error(None)
    # a particle outside the sorting range sets the error bit, as
    # fipi_cells_build reports FIPI_ERR_RANGE
    p = config_params(32, 10, 2.0, 100)
    pos = fipi_amd.init_rsa(10, 2.0, 0.5, (p.lx, p.ly, p.lz), 3)
    pos[4, 1] = -3.0
    _, _, err = _noalloc(p, pos)
    assert err == 1
    # and the next call starts clean
    pos[4, 1] = 3.0
    s, i, err = _noalloc(p, pos)
    assert err == 0
    r = _with_ctx(p, pos)
    assert np.array_equal(s, r[0]) and np.array_equal(i, r[1])
