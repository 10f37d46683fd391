"""FP64 and FP32-mixed over 1000 steps against oracle fixtures
(tests/golden/long_*.npz, written by tests/golden/make_long_horizon.py in
this repository).

The fixture holds, every 50 (128^3) or 100 (256^3) steps, phi and u at
seeded grid points and their global L2 norms from the oracle, the
positions of the spheres at a few checkpoints, and the divergence of two
oracle runs that differ only in their FFT factor order (radix 4 + 2
against radix 2).  Up to the step where the two oracle runs' positions
still agree to 1e-12, phi and u must agree at 1e-10 and positions at
1e-10.  Beyond it (a sphere on a vanishing grad c turns its capillary
force with the round-off, and any two FP64 implementations part there;
where such a sphere moves far enough, its force reaches phi and u) the GPU
may differ from the oracle by no more than 100 times what the two oracle
runs differ by (plus the bars above).  The FP32-mixed path is held to the
FP32-mixed tolerance against the same fixtures.
"""
import os

import numpy as np
import pytest

import fipi_amd

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture(n):
    path = os.path.join(HERE, "golden", f"long_{n}.npz")
    if not os.path.exists(path):
        pytest.skip(f"fixture {path} not generated")
    return np.load(path)


@pytest.mark.parametrize("n", [128, 256])
def test_fp64_1000_steps_against_oracle_fixture(gpu, n):
    fx = _fixture(n)
    npart, steps = int(fx["particles"]), int(fx["steps"])
    # the workload of a run of run_steps (its contact ramp spans 0.2 of it),
    # checked through the steps the fixture holds
    run_steps = int(fx["run_steps"]) if "run_steps" in fx.files else steps
    print(f"long horizon {n}^3 + {npart}, {steps} of {run_steps} steps", flush=True)
    p = fipi_amd.config_params(n, npart, 2.0, run_steps)
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, 1)
    pos = fipi_amd.init_rsa(npart, 2.0, 0.5, (p.lx, p.ly, p.lz), 2)
    assert pos.shape[0] == npart
    cth = np.cos(np.deg2rad(fipi_amd.init_uniform(npart, 60.0, 120.0, 3)))
    idx, pidx = fx["sample_idx"], fx["pos_idx"]
    ab = fx["ab_pos_rel"]
    # last step at which the two oracle runs still agree to 1e-12
    apart = np.nonzero(ab > 1e-12)[0]
    horizon = int(apart[0]) if apart.size else steps
    L = np.array([p.lx, p.ly, p.lz])
    pos_at = {int(t): k for k, t in enumerate(fx["pos_steps"])}
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0, pos, None, cth)
    t = 0
    for k, tc in enumerate(fx["ck_steps"]):
        sim.step(int(tc) - t)
        t = int(tc)
        phi = sim.phi().reshape(-1)
        u = sim.velocity().reshape(3, -1)
        rphi, ru = fx["ck_phi"][k], fx["ck_u"][k]
        e_phi = np.linalg.norm(phi[idx] - rphi) / np.linalg.norm(rphi)
        e_u = np.linalg.norm(u[:, idx] - ru) / np.linalg.norm(ru)
        n_phi = abs(np.linalg.norm(phi) / fx["ck_phi_norm"][k] - 1)
        n_u = abs(np.linalg.norm(u) / fx["ck_u_norm"][k] - 1)
        # the two oracle runs' own difference in phi and u at this checkpoint
        ab_phi, ab_u = float(fx["ck_ab_phi"][k]), float(fx["ck_ab_u"][k])
        msg = (f"t={t}: phi {e_phi:.2e} (norm {n_phi:.1e}) u {e_u:.2e} (norm {n_u:.1e}) "
               f"[two oracles: phi {ab_phi:.1e} u {ab_u:.1e}]")
        if t in pos_at:
            gp = sim.particles()["p"][pidx]
            rp = fx["pos"][pos_at[t]]
            d = gp - rp
            d -= L * np.round(d / L)
            e_p = np.linalg.norm(d) / np.linalg.norm(rp)
            msg += f" positions {e_p:.2e} (two oracles {ab[t - 1]:.2e})"
            if t <= horizon:
                assert e_p < 1e-10, msg
            else:
                assert e_p < 100 * ab[t - 1] + 1e-10, msg
        print(msg, flush=True)
        if t <= horizon:
            assert e_phi < 1e-10 and e_u < 1e-10, msg
            assert n_phi < 1e-12 and n_u < 1e-12, msg
        else:
            # past the horizon a parted sphere's force reaches the flow
            assert e_phi < 1e-10 + 100 * ab_phi and e_u < 1e-10 + 100 * ab_u, msg
            assert n_phi < 1e-12 + 100 * ab_phi and n_u < 1e-12 + 100 * ab_u, msg
    sim.close()
    print(f"oracle orders agree to 1e-12 through step {horizon}", flush=True)


@pytest.mark.parametrize("n", [128, 256])
def test_fp32_mixed_against_oracle_fixture(gpu, n):
    """FP32-mixed (FFTs, k-space update, spread and interpolation in FP32)
    against the FP64 oracle fixture over the same steps: phi at the sample
    points and its global norm within 1e-5 relative (the FP32-mixed
    tolerance), u within 1e-3 at the sample points (the tolerance names no
    bar for u; FP32 transforms of the small flow field leave ~1e-4), and no
    sphere more than 1e-3 dx from the oracle's position at the position
    checkpoints, plus 100 times the largest displacement between the two
    oracle runs there (a sphere that two FP64 runs already part on)."""
    fx = _fixture(n)
    npart, steps = int(fx["particles"]), int(fx["steps"])
    run_steps = int(fx["run_steps"]) if "run_steps" in fx.files else steps
    print(f"FP32-mixed against the oracle fixture {n}^3 + {npart}, {steps} of {run_steps} steps", flush=True)
    p = fipi_amd.config_params(n, npart, 2.0, run_steps, precision=fipi_amd.FP32_MIXED)
    phi0 = fipi_amd.init_phi((n, n, n), 0.0, 0.05, 1)
    pos = fipi_amd.init_rsa(npart, 2.0, 0.5, (p.lx, p.ly, p.lz), 2)
    cth = np.cos(np.deg2rad(fipi_amd.init_uniform(npart, 60.0, 120.0, 3)))
    idx, pidx = fx["sample_idx"], fx["pos_idx"]
    L = np.array([p.lx, p.ly, p.lz])
    pos_at = {int(t): k for k, t in enumerate(fx["pos_steps"])}
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0, pos, None, cth)
    t = 0
    for k, tc in enumerate(fx["ck_steps"]):
        sim.step(int(tc) - t)
        t = int(tc)
        phi = sim.phi().reshape(-1)
        u = sim.velocity().reshape(3, -1)
        rphi, ru = fx["ck_phi"][k], fx["ck_u"][k]
        e_phi = np.linalg.norm(phi[idx] - rphi) / np.linalg.norm(rphi)
        e_u = np.linalg.norm(u[:, idx] - ru) / np.linalg.norm(ru)
        n_phi = abs(np.linalg.norm(phi) / fx["ck_phi_norm"][k] - 1)
        msg = f"t={t}: phi {e_phi:.2e} (norm {n_phi:.1e}) u {e_u:.2e}"
        if t in pos_at:
            d = sim.particles()["p"][pidx] - fx["pos"][pos_at[t]]
            d -= L * np.round(d / L)
            dmax = float(np.abs(d).max())
            ab_max = float(fx["ab_max_disp"][t - 1])
            msg += f" max |dp| {dmax:.2e} dx (two oracles {ab_max:.2e})"
            assert dmax < 1e-3 + 100 * ab_max, msg
        print(msg, flush=True)
        assert e_phi < 1e-5 and n_phi < 1e-5 and e_u < 1e-3, msg
    sim.close()

