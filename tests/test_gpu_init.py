"""Seeded initial state generated on the GPU (fipi_init_*_device,
fipi_set_state_device) against the host generators: the same phi0,
contact angles and random-sequential-addition positions bit for bit --
the headline's 100,000 spheres in 256^3, the 128^3 density, a box that
jams before n (max_attempts), overlapping spheres (gap < 0), a box with
fewer than 3 RSA cells a side and a non-cubic box -- and a state set from
device arrays steps exactly as the one set from the host."""
import time

import numpy as np
import pytest

import fipi_amd
from fipi_amd import FP32_MIXED, FP64

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(64, 64, 64), (256, 256, 256), (48, 80, 96)])
def test_phi_and_uniform(gpu, shape):
    print(f"init phi {shape}", flush=True)
    d = fipi_amd.init_phi_device(shape, 0.1, 0.05, 17).cpu().numpy()
    assert np.array_equal(d, fipi_amd.init_phi(shape, 0.1, 0.05, 17))
    n = int(np.prod(shape)) // 7
    assert np.array_equal(fipi_amd.init_uniform_device(n, 60.0, 120.0, 3).cpu().numpy(),
                          fipi_amd.init_uniform(n, 60.0, 120.0, 3))


RSA_CASES = [
    # n, r, gap, box, seed, max_attempts
    (100000, 2.0, 0.5, (256.0, 256.0, 256.0), 2, 100000),   # the headline workload
    (12500, 2.0, 0.5, (128.0, 128.0, 128.0), 2, 100000),    # its density at 128^3
    (5000, 2.5, 0.5, (64.0, 64.0, 64.0), 7, 2000),          # jams before n
    (3000, 2.0, -1.0, (48.0, 64.0, 80.0), 11, 100000),      # overlapping, non-cubic
    (30, 1.5, 0.2, (8.0, 7.0, 9.0), 5, 500),                # fewer than 3 cells a side
]


@pytest.mark.parametrize("case", RSA_CASES)
def test_rsa_matches_host(gpu, case):
    n, r, gap, box, seed, att = case
    print(f"init rsa n={n} r={r} gap={gap} box={box}", flush=True)
    t0 = time.time()
    host = fipi_amd.init_rsa(n, r, gap, box, seed, att)
    t1 = time.time()
    dev = fipi_amd.init_rsa_device(n, r, gap, box, seed, att).cpu().numpy()
    t2 = time.time()
    print(f"placed {host.shape[0]} of {n}: host {t1 - t0:.3f} s, device {t2 - t1:.3f} s", flush=True)
    assert dev.shape == host.shape
    assert np.array_equal(dev, host)


@pytest.mark.parametrize("prec", [FP64, FP32_MIXED])
def test_state_from_device_arrays_steps_as_from_host(gpu, prec):
    import torch
    print(f"set_state_device prec={prec}", flush=True)
    n, npart = 64, 1500
    p = fipi_amd.config_params(n, npart, 2.0, 50, precision=prec)
    phi_d = fipi_amd.init_phi_device((n, n, n), 0.0, 0.05, 1)
    pos_d = fipi_amd.init_rsa_device(npart, 2.0, 0.5, (p.lx, p.ly, p.lz), 2)
    assert pos_d.shape[0] == npart
    cth_d = torch.cos(torch.deg2rad(fipi_amd.init_uniform_device(npart, 60.0, 120.0, 3)))
    a = fipi_amd.Simulation(p)
    a.set_state_device(phi_d, pos_d, None, cth_d)
    a.step(6)
    b = fipi_amd.Simulation(p)
    b.set_state(phi_d.cpu().numpy(), pos_d.cpu().numpy(), None, cth_d.cpu().numpy())
    b.step(6)
    assert np.array_equal(a.phi(), b.phi())
    assert np.array_equal(a.particles()["p"], b.particles()["p"])
    assert a.diagnostics()["mean_phi"] == b.diagnostics()["mean_phi"]
    a.close()
    b.close()


def test_device_init_errors(gpu):
    with pytest.raises(fipi_amd.FipiError):
        fipi_amd.init_rsa_device(10, -1.0, 0.5, (16.0, 16.0, 16.0), 1)
    with pytest.raises(fipi_amd.FipiError):
        fipi_amd.init_rsa_device(10, 1.0, -2.5, (16.0, 16.0, 16.0), 1)
