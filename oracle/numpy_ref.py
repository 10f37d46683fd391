"""Independent NumPy restatement of the reference time step -- TEST
INFRASTRUCTURE ONLY.

Written separately from oracle/fipi_oracle.c (vectorised, numpy.fft for
FFTW, Python loops for particles) so that the two restatements cross-check
each other on tiny grids.  Follows src/main.cpp:75-222 statement by
statement; every helper cites the reference lines it restates.  Only for
small cases (<= 16^3, tens of particles).
"""
from __future__ import annotations

import math

import numpy as np


class NumpyRef:
    def __init__(self, p, phi0, pos=None, radius=None, cos_theta=None):
        self.p = p
        nx, ny, nz = p.nx, p.ny, p.nz
        self.shape = (nz, ny, nx)
        self.G = nx * ny * nz
        # field.cpp:108-115 wave numbers (index i -> i or i-N, times 2*pi/L)
        def kv(n, L):
            return np.array([(i if i <= n // 2 else i - n) * 2.0 * p.pi / L for i in range(n)])
        kx, ky, kz = kv(nx, p.lx), kv(ny, p.ly), kv(nz, p.lz)
        self.K = [kx[None, None, :] + 0 * ky[None, :, None] + 0 * kz[:, None, None],
                  0 * kx[None, None, :] + ky[None, :, None] + 0 * kz[:, None, None],
                  0 * kx[None, None, :] + 0 * ky[None, :, None] + kz[:, None, None]]
        self.K2 = (self.K[0] ** 2 + self.K[1] ** 2) + self.K[2] ** 2
        # Nyquist masks of grad (field.cpp:310-318)
        ii = np.arange(nx)[None, None, :] + 0 * np.zeros(self.shape, int)
        jj = np.arange(ny)[None, :, None] + 0 * np.zeros(self.shape, int)
        kk = np.arange(nz)[:, None, None] + 0 * np.zeros(self.shape, int)
        self.nyq = [ii == nx // 2, jj == ny // 2, kk == nz // 2]
        self.c = np.asarray(phi0, dtype=np.float64).reshape(self.shape).astype(np.complex128)
        self.u = [np.zeros(self.shape, np.complex128) for _ in range(3)]
        self.grad = [np.zeros(self.shape, np.complex128) for _ in range(3)]
        self.t = 0
        self.n = p.n_particles if p.switch_particle else 0
        n = self.n
        self.pos = np.array(pos, dtype=np.float64).copy() if n else np.zeros((0, 3))
        self.pos_old = self.pos.copy()
        self.pu = np.zeros((n, 3))
        self.pu_old = np.zeros((n, 3))
        self.f_pp = np.zeros((n, 3))
        self.f_pi = np.zeros((n, 3))
        self.rad = np.full(n, p.r_p) if radius is None else np.asarray(radius, float)
        self.cth = np.zeros(n) if cos_theta is None else np.asarray(cos_theta, float)
        self.it_mark = -1
        # cluster.cpp:26-32
        self.nc, self.dc = [], []
        for L in (p.lx, p.ly, p.lz):
            m = int(math.floor(L / p.r_c + 2 * p.margin / p.r_c))
            self.nc.append(m)
            self.dc.append(math.fmod(L + 2 * p.margin, p.r_c) / m + p.r_c)

    # field::fft / field::ifft (field.cpp:438-455) with numpy's pocketfft
    @staticmethod
    def fft(a):
        return np.fft.fftn(a)

    @staticmethod
    def ifft(a):
        return np.fft.ifftn(a)

    def f_of(self, c):
        p = self.p
        return (p.lam / (p.epsilon * p.epsilon)) * (np.real(c) ** 3 - np.real(c)) + 0j

    def grad_of(self, chat):
        return [np.where(self.nyq[d], 0.0, 1j * self.K[d] * chat) for d in range(3)]

    # ---------------- particles ------------------------------------------
    def _interp(self, v, x):
        # field.cpp:223-248
        p = self.p
        nx, ny, nz = p.nx, p.ny, p.nz
        dx, dy, dz = p.lx / nx, p.ly / ny, p.lz / nz
        v = np.real(v)
        i, j, k = int(math.floor(x[0] / dx)), int(math.floor(x[1] / dy)), int(math.floor(x[2] / dz))
        m, n, o = x[0] / dx - math.floor(x[0] / dx), x[1] / dy - math.floor(x[1] / dy), x[2] / dz - math.floor(x[2] / dz)
        i1, j1, k1 = (i + 1) % nx, (j + 1) % ny, (k + 1) % nz
        val = 0.0
        for (a, wa) in ((i, 1 - m), (i1, m)):
            for (b, wb) in ((j, 1 - n), (j1, n)):
                for (c, wc) in ((k, 1 - o), (k1, o)):
                    val += wa * wb * wc * v[c, b, a]
        return val

    def _cell(self, x):
        return [int(math.floor(x[a] / self.dc[a] + self.p.margin / self.dc[a])) for a in range(3)]

    def _f_pi(self, phi, g, r, ct):
        # element.cpp:75-97 (+ wetting offset d0 = r cos(theta), zero at 90 deg)
        p = self.p
        if abs(phi) <= p.tanh2:
            gn = math.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
            if gn != 0.0:
                d = 0.5 * math.log((1 + phi) / (1 - phi)) * math.sqrt(2.0) * p.epsilon - r * ct
                fm = -p.a_pi * p.pi * d * p.sigma if abs(d) < r else 0.0
                return np.array(g) / gn * fm
        return np.zeros(3)

    def _forces(self, tp1):
        # cluster.cpp:174-275, brute force over pairs within the 27-cell
        # neighbourhood; order: reference scatter order is not reproduced
        # here (sums are compared to rounding)
        p = self.p
        n = self.n
        f = np.zeros((n, 3))
        ramp = 0.0 if tp1 < 0 else (tp1 / p.ramp_t if tp1 < p.ramp_t else 1.0)
        L = np.array([p.lx, p.ly, p.lz])
        th = np.array([self.dc[a] * (self.nc[a] - 2) for a in range(3)])
        cells = [self._cell(x) for x in self.pos]
        for i in range(n):
            for j in range(i + 1, n):
                near = all(((cells[i][a] - cells[j][a]) % self.nc[a]) in (0, 1, self.nc[a] - 1) for a in range(3))
                if not near:
                    continue
                d = self.pos[i] - self.pos[j]
                d = np.where(d > th, d - L, np.where(d < -th, d + L, d))
                nr = math.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
                if nr < p.r_eq:
                    F = (p.r_eq - nr) * ramp * p.k_n * (d / nr)
                    f[i] += F
                    f[j] -= F
        return f

    def _stage(self, stage, tp1):
        p = self.p
        for a in range(self.n):
            x = self.pos[a]
            phi = self._interp(self.c, x)
            g = [self._interp(self.grad[d], x) for d in range(3)]
            self.f_pi[a] = self._f_pi(phi, g, self.rad[a], self.cth[a])
        self.f_pp = self._forces(tp1)
        v = (self.f_pi + self.f_pp) / 6.0 / p.pi / p.mu / self.rad[:, None]
        if stage == 1:
            self.pu_old = v.copy()
            self.pos_old = self.pos.copy()
            self.pos = self.pos + p.dt * v
        else:
            self.pu = v.copy()
            self.pos = self.pos_old + 0.5 * p.dt * (self.pu_old + v)
        L = np.array([p.lx, p.ly, p.lz])
        self.pos = np.fmod(self.pos + L, L)

    def _spread(self):
        # cluster.cpp:344-410
        p = self.p
        nx, ny, nz = p.nx, p.ny, p.nz
        dx, dy, dz = p.lx / nx, p.ly / ny, p.lz / nz
        out = np.zeros((3,) + self.shape)
        for a in range(self.n):
            f = self.f_pp[a] / dx / dy / dz
            x = self.pos[a]
            i, j, k = int(math.floor(x[0] / dx)), int(math.floor(x[1] / dy)), int(math.floor(x[2] / dz))
            o, q, r = x[0] / dx - math.floor(x[0] / dx), x[1] / dy - math.floor(x[1] / dy), x[2] / dz - math.floor(x[2] / dz)
            for (aa, wa) in ((i, 1 - o), ((i + 1) % nx, o)):
                for (bb, wb) in ((j, 1 - q), ((j + 1) % ny, q)):
                    for (cc, wc) in ((k, 1 - r), ((k + 1) % nz, r)):
                        out[:, cc, bb, aa] += f * wa * wb * wc
        return out

    # ---------------- one reference step (main.cpp:75-222) ---------------
    def step(self, nsteps=1):
        p = self.p
        for _ in range(nsteps):
            t = self.t
            # CH, main.cpp:81-140
            f = self.fft(self.f_of(self.c))
            chat = self.fft(self.c)
            conv = np.zeros(self.shape, np.complex128)
            if p.switch_ns:
                self.grad = [self.ifft(g) for g in self.grad_of(chat)]
                conv = self.fft(sum(np.real(self.u[d]) * np.real(self.grad[d]) for d in range(3)) + 0j)
            K2, K4 = self.K2, self.K2 ** 2
            if t == 0:
                self.c_old, self.f_old, self.conv_old = chat, f, conv
                cn = (chat - p.dt * p.mobility * K2 * f - p.dt * conv) / (1.0 + p.dt * p.mobility * p.lam * K4)
            else:
                cn = (4.0 * chat - self.c_old - 4.0 * p.dt * p.mobility * K2 * f
                      + 2.0 * p.dt * p.mobility * K2 * self.f_old - 4.0 * p.dt * conv
                      + 2.0 * p.dt * self.conv_old) / (3.0 + 2.0 * p.dt * p.mobility * p.lam * K4)
                self.c_old, self.f_old, self.conv_old = chat, f, conv
            self.c = self.ifft(cn)
            # Stokes, main.cpp:149-179
            if p.switch_ns:
                f = self.fft(self.f_of(self.c))
                chat = self.fft(self.c)
                mu = np.real(self.ifft(p.lam * K2 * chat + f))
                self.grad = [self.ifft(g) for g in self.grad_of(chat)]
                self.c = self.ifft(chat)
                F = [mu * np.real(self.grad[d]) for d in range(3)]
                if self.n:
                    s = self._spread()
                    F = [F[d] + s[d] for d in range(3)]
                cr = np.real(self.c)
                b = (1.0 + cr) / 2.0 * p.rho1 + (1.0 - cr) / 2.0 * p.rho2
                F[2] = F[2] + b * p.gravity
                Fh = [self.fft(F[d] + 0j) for d in range(3)]
                kdotF = sum(self.K[d] * Fh[d] for d in range(3))
                with np.errstate(divide="ignore", invalid="ignore"):
                    uh = [np.where(K2 != 0, (Fh[d] - self.K[d] * kdotF / np.where(K2 != 0, K2, 1.0))
                                   / np.where(K2 != 0, p.mu * K2, 1.0), 0.0) for d in range(3)]
                self.u = [self.ifft(uh[d]) for d in range(3)]
            # particles, main.cpp:186-219
            if self.n:
                chat = self.fft(self.c)
                self.grad = [self.ifft(g) for g in self.grad_of(chat)]
                self.c = self.ifft(chat)
                self._stage(1, t + 1)
                self._stage(2, t + 1)
            self.t += 1

    def phi(self):
        return np.real(self.c)

    def velocity(self):
        return np.array([np.real(u) for u in self.u])
