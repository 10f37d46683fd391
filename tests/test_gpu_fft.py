"""GPU 3-D real transforms (libfipi fipi_fft_r2c / fipi_fft_c2r) against
numpy.fft (pocketfft), the stand-in for FFTW's fftw_execute_dft
(src/field.cpp:438-455)."""
import numpy as np
import pytest

import fipi_amd
from fipi_amd import FP32_MIXED, FP64, config_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [16, 32, 48, 64, 80, 96, 128, 256, (48, 80, 96), (96, 32, 80)])
@pytest.mark.parametrize("prec", [FP64, FP32_MIXED])
def test_r2c_c2r_vs_numpy(gpu, n, prec):
    """Radix-2/4/8 lengths and the radix-3/5/6 middle stages of 48, 80 and
    96, cubic and mixed (nx, ny, nz)."""
    import torch
    print(f"fft n={n} prec={prec}", flush=True)
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    p = config_params(nx, 0, 2.0, 10, precision=prec).replace(ny=ny, nz=nz, ly=float(ny), lz=float(nz))
    sim = fipi_amd.Simulation(p)
    rng = np.random.default_rng(nx * 7 + ny * 3 + nz)
    x = rng.standard_normal((nz, ny, nx))
    dt = torch.float64 if prec == FP64 else torch.float32
    xt = torch.tensor(x, dtype=dt, device=gpu)
    X = sim.fft_r2c(xt)
    Xc = X.cpu().numpy().astype(np.float64)
    Xc = Xc[..., 0] + 1j * Xc[..., 1]
    ref = np.fft.rfftn(x.astype(np.float32).astype(np.float64) if prec else x)
    err = np.linalg.norm(Xc - ref) / np.linalg.norm(ref)
    tol = 1e-14 if prec == FP64 else 2e-6
    assert err < tol, err
    y = sim.fft_c2r(X).cpu().numpy().astype(np.float64)
    xr = x.astype(np.float32).astype(np.float64) if prec else x
    err2 = np.linalg.norm(y - xr) / np.linalg.norm(xr)
    assert err2 < tol, err2
    sim.close()


def test_single_modes_fp64(gpu):
    """Every single Fourier mode of a 16^3 grid maps to one bin (all index
    combinations, including Nyquist planes)."""
    import torch
    n = 16
    p = config_params(n, 0, 2.0, 10)
    sim = fipi_amd.Simulation(p)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    for (a, b, c) in [(0, 0, 0), (1, 0, 0), (0, 3, 0), (0, 0, 5), (8, 0, 0), (0, 8, 8), (3, 7, 2), (8, 8, 8)]:
        f = np.cos(2 * np.pi * (a * x + b * y + c * z) / n)
        X = sim.fft_r2c(torch.tensor(f, dtype=torch.float64, device=gpu)).cpu().numpy()
        Xc = X[..., 0] + 1j * X[..., 1]
        ref = np.fft.rfftn(f)
        assert np.abs(Xc - ref).max() < 1e-11 * n ** 3
    sim.close()


@pytest.mark.parametrize("ny", [64, 128, 256])
@pytest.mark.parametrize("prec", [FP64, FP32_MIXED])
def test_y_plans_vs_numpy(gpu, ny, prec):
    """Both FFT plans of the y passes that a tuner may pick (256: 8.4.8 and
    4.4.4.4 in FP64, 16.16 and 8.4.8 in FP32; 128: 8.2.8 and 4.8.4; 64: 8.8
    and 4.4.4) against numpy, on a grid that is long in y."""
    import torch
    print(f"y plans ny={ny} prec={prec}", flush=True)
    nx, nz = 32, 16
    p = config_params(nx, 0, 2.0, 10, precision=prec).replace(ny=ny, nz=nz, ly=float(ny), lz=float(nz))
    sim = fipi_amd.Simulation(p)
    default = sim.tuning()["y_values_per_thread"]
    plans = ({8, 4} if prec == FP64 else {8, 16}) if ny == 256 else {8, 4}
    assert default in plans
    rng = np.random.default_rng(ny + prec)
    x = rng.standard_normal((nz, ny, nx))
    dt = torch.float64 if prec == FP64 else torch.float32
    xr = x.astype(np.float32).astype(np.float64) if prec else x
    ref = np.fft.rfftn(xr)
    tol = 1e-14 if prec == FP64 else 2e-6
    for e in sorted(plans):
        t = sim.tuning()
        t["y_values_per_thread"] = e
        sim.set_tuning(t)
        assert sim.tuning()["y_values_per_thread"] == e
        X = sim.fft_r2c(torch.tensor(x, dtype=dt, device=gpu))
        Xc = X.cpu().numpy().astype(np.float64)
        err = np.linalg.norm(Xc[..., 0] + 1j * Xc[..., 1] - ref) / np.linalg.norm(ref)
        assert err < tol, (e, err)
        y = sim.fft_c2r(X).cpu().numpy().astype(np.float64)
        assert np.linalg.norm(y - xr) / np.linalg.norm(xr) < tol, e
    with pytest.raises(fipi_amd.FipiError):
        t = sim.tuning()
        t["y_values_per_thread"] = 5
        sim.set_tuning(t)
    sim.close()
