"""Runtime parameters of the FIPI step (the reference's compile-time
constants, src/head.hpp:27-74 and src/parameter.hpp:18-20) and the presets
the tests and the benchmark use.
"""
from __future__ import annotations

import dataclasses
import math

FP64 = 0
FP32_MIXED = 1

# src/head.hpp:27
REFERENCE_PI = 3.1415926


@dataclasses.dataclass
class Params:
    nx: int = 64
    ny: int = 64
    nz: int = 64
    n_particles: int = 0
    lx: float = 64.0
    ly: float = 64.0
    lz: float = 64.0
    pi: float = REFERENCE_PI
    dt: float = 0.01
    ramp_t: float = 40.0
    epsilon: float = 1.0
    lam: float = 1.0
    mobility: float = 1.0
    mu: float = 1.0
    sigma: float = 0.0
    a_pi: float = 2.0
    k_n: float = 0.0
    r_p: float = 2.0
    r_eq: float = 4.0
    r_c: float = 6.0
    margin: float = 0.0
    rho1: float = 1e-3
    rho2: float = 1e-3
    gravity: float = -10.0
    tanh2: float = math.tanh(2.0)
    switch_particle: int = 1
    switch_ch: int = 1
    switch_ns: int = 1
    precision: int = FP64
    interp_u: int = 0
    kt: float = 0.0
    noise_seed: int = 0

    @property
    def G(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def dx(self) -> float:
        return self.lx / self.nx

    def replace(self, **kw) -> "Params":
        return dataclasses.replace(self, **kw)

    def as_dict(self) -> dict:
        return dataclasses.asdict(self)

    def cell_grid(self):
        """Contact cell grid, src/cluster.cpp:26-32."""
        n, d = [], []
        for L in (self.lx, self.ly, self.lz):
            nc = int(math.floor(L / self.r_c + 2 * self.margin / self.r_c))
            n.append(nc)
            d.append(math.fmod(L + 2 * self.margin, self.r_c) / nc + self.r_c)
        return n, d


def reference_params(n_particles: int = 20000, t_end: int = 20000) -> Params:
    """The reference's own constants, src/head.hpp:27-74 (64^3, L = 2 pi,
    dt = 1e-3, sigma = 0.07, eps = 3 dx/sqrt(2)/2, ...), evaluated with the
    same expressions."""
    pi = REFERENCE_PI
    nx = ny = nz = 64
    lx = ly = lz = 2 * pi
    dx = lx / nx
    sigma = 0.07
    epsilon = 3.0 * dx / pow(2, 0.5) / 2.0
    lam = sigma * epsilon * 3.0 / 2.0 / pow(2, 0.5)
    r_drop = 0.4 * lx
    gravity = -1e1
    rho1 = 1e-3
    Bo = 0.0
    rho2 = rho1 - Bo * sigma / gravity / r_drop / r_drop
    mu = 1e-3
    mobility = pow(0.2 * epsilon, 2) / mu
    r_p = 0.03 * r_drop
    return Params(
        nx=nx, ny=ny, nz=nz, n_particles=n_particles, lx=lx, ly=ly, lz=lz, pi=pi,
        dt=1e-3, ramp_t=0.2 * t_end, epsilon=epsilon, lam=lam, mobility=mobility, mu=mu,
        sigma=sigma, a_pi=2.0, k_n=0.5 * sigma, r_p=r_p, r_eq=5 * r_p, r_c=5 * r_p,
        margin=0.0, rho1=rho1, rho2=rho2, gravity=gravity, tanh2=math.tanh(2.0),
    )


def config_params(n: int, n_particles: int, r_p: float, steps: int,
                  precision: int = FP64, cell: float = 6.0) -> Params:
    """Benchmark configurations: periodic n^3 box with dx = 1, eps = 1,
    M = 1, kappa (lambda) = 1, dt = 0.01.  The reference leaves the fluid
    viscosity and the particle constants unspecified for these units; they
    are fixed here as: eta = 1; sigma = 2 sqrt(2) lambda / (3 eps) (the
    inverse of head.hpp:47); A = 2 (head.hpp:57); k_n = sigma/2
    (parameter.hpp:18); contact cutoff r_eq = 2 r_p (touching spheres; must
    not exceed the cell size); contact cells of `cell` dx; ramp over
    0.2 * steps (cluster.cpp:249); no buoyancy contrast (Bo = 0)."""
    eps = 1.0
    lam = 1.0
    sigma = 2.0 * math.sqrt(2.0) * lam / (3.0 * eps)
    return Params(
        nx=n, ny=n, nz=n, n_particles=n_particles, lx=float(n), ly=float(n), lz=float(n),
        pi=REFERENCE_PI, dt=0.01, ramp_t=0.2 * steps, epsilon=eps, lam=lam, mobility=1.0,
        mu=1.0, sigma=sigma, a_pi=2.0, k_n=0.5 * sigma, r_p=r_p, r_eq=2.0 * r_p, r_c=cell,
        margin=0.0, rho1=1e-3, rho2=1e-3, gravity=-10.0, tanh2=math.tanh(2.0),
        switch_particle=1 if n_particles > 0 else 0, precision=precision,
    )
