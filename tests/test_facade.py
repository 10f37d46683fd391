"""The C++ facade (fipi_amd/facade/fipi.hpp) keeps the reference's class
interfaces (field, cluster, element, list, vector, output_time,
convert_to_vtk) over the C ABI.

* fipi_main has the shape of src/main.cpp with the loop body replaced by
  the fused GPU step; its diagnostics equal the Python binding's.
* fipi_loop writes main.cpp's loop body (main.cpp:75-222) statement by
  statement through the facade's device fields and cluster methods and
  checks 5 steps against the fused step at 1e-12 (phi, u, positions).
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FACADE = os.path.join(ROOT, "fipi_amd", "facade")


def test_facade_builds():
    subprocess.run(["make", "-s", "-C", FACADE], check=True)
    assert os.path.exists(os.path.join(FACADE, "fipi_main"))
    assert os.path.exists(os.path.join(FACADE, "fipi_loop"))


@pytest.mark.gpu
def test_facade_driver_matches_python_mirror(gpu, tmp_path):
    import fipi_amd
    exe = os.path.join(FACADE, "fipi_main")
    out = subprocess.run([exe, "32", "100", "20"], capture_output=True, text=True, timeout=120,
                         cwd=tmp_path)
    assert out.returncode == 0, out.stderr
    m = re.search(r"steps 20 .* mean phi (\S+) free energy (\S+)", out.stdout)
    assert m, out.stdout
    p = fipi_amd.config_params(32, 100, 2.5, 20)
    phi0 = fipi_amd.init_phi((32, 32, 32), 0.0, 0.05, 1)
    pos = fipi_amd.init_rsa(100, 2.5, 0.5, (32.0, 32.0, 32.0), 2)
    sim = fipi_amd.Simulation(p)
    sim.set_state(phi0, pos)
    sim.step(20)
    d = sim.diagnostics()
    sim.close()
    assert float(m.group(1)) == d["mean_phi"]
    assert float(m.group(2)) == d["free_energy"]


@pytest.mark.gpu
def test_facade_loop_body_matches_fused_step(gpu, tmp_path):
    print("facade loop body vs fused step, 32^3 + 300 spheres, 5 steps", flush=True)
    exe = os.path.join(FACADE, "fipi_loop")
    out = subprocess.run([exe, "32", "300", "5", "20"], capture_output=True, text=True, timeout=300,
                         cwd=tmp_path)
    print(out.stdout, flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    m = re.search(r"phi (\S+) u (\S+) positions (\S+) tau (\S+) vs (\S+)", out.stdout)
    assert m, out.stdout
    e_phi, e_u, e_pos, tau = (float(m.group(i)) for i in range(1, 5))
    assert e_phi < 1e-12 and e_u < 1e-12 and e_pos < 1e-12
    assert tau != 0.0  # the spheres touch: the packing energy is exercised
