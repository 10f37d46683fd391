"""Seeded test cases shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np

import fipi_amd
from fipi_amd import FP64, config_params


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def make_case(n=32, n_particles=150, r_p=2.0, steps=100, seed=1, precision=FP64,
              ns=1, mean=0.0, amp=0.05, gap=0.5, cell=6.0, cap=False):
    """Periodic n^3 box in the benchmark units (dx = 1, eps = 1, M = 1,
    kappa = 1, dt = 0.01), phi0 = mean + U(-amp, amp), spheres by seeded
    random sequential addition.  cap: a request beyond what the addition
    places is capped at the count it placed.  Returns (params, phi0, pos)."""
    p = config_params(n, n_particles, r_p, steps, precision=precision, cell=cell)
    p = p.replace(switch_ns=ns)
    phi0 = fipi_amd.init_phi((n, n, n), mean, amp, seed)
    pos = None
    if n_particles > 0:
        pos = fipi_amd.init_rsa(n_particles, r_p, gap, (p.lx, p.ly, p.lz), seed + 1000)
        if cap:
            p = p.replace(n_particles=pos.shape[0])
        assert pos.shape[0] == p.n_particles, f"RSA placed {pos.shape[0]} of {n_particles}"
    return p, phi0, pos
