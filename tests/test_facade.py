"""The C++ facade (fipi_amd/facade/fipi.hpp) keeps the reference's class
interfaces (field, cluster, vector, output_time, convert_to_vtk) over the C
ABI; its main() has the shape of src/main.cpp with the loop body replaced
by the GPU step."""
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
