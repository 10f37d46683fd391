// Field kernels of the spectral Cahn-Hilliard / Stokes step.
//
// Layouts (x fastest, as in the reference: it = i + Nx*(j + Ny*k),
// src/field.cpp:88-92):
//   real field     R[z][y][x]                      NX reals per line
//   half spectrum  H[z][y][kx], kx in [0, NX/2]    NXH = NX/2+1 complex per line
// The same half-spectrum layout holds data after the x transform (X), after
// x and y (XY) and the full spectrum (K), so every pass and the k-space
// history index one array shape.
//
// Pass structure (each pass reads and writes every array once):
//   x pass   two real lines packed as one complex line (a + i b), one
//            complex transform, split into the two half spectra (R2C), or
//            the reverse (C2R).  32 real lines per workgroup.
//   y pass   16 neighbouring kx columns x all y at one z (256 B runs).
//   z hub    16 kx columns x all z at one y: forward z transform, the
//            fused k-space update, inverse z transform, all in registers.
// Wave numbers are read from per-axis tables computed on the host with the
// reference expression i*2.0*pi/L (src/field.cpp:108-115).
#pragma once
#include "fft_core.hpp"

namespace fipi {

// nxh = nx/2 + 1 spectral columns; nxp >= nxh is the row pitch of every
// half-spectrum array, a multiple of 128 bytes so that the 16-column tiles
// of the y and z passes start on cache-line boundaries.
struct Dims {
    int nx, ny, nz, nxh, nxp;
};

constexpr int TILE = 16;  // kx columns per y/z workgroup; line pairs per x workgroup

__device__ __forceinline__ size_t hidx(const Dims &d, int kx, int y, int z)
{
    return (size_t)kx + (size_t)d.nxp * ((size_t)y + (size_t)d.ny * z);
}

// ---------------------------------------------------------------------------
// x pass helpers
// ---------------------------------------------------------------------------

// Load the packed complex line Z = A + iB of real lines (rl, rl+1) in
// distribution D(N1): thread t < N1 gets Z[t + N1*m].  A = in[rl], B = in[rl+1]
// (half spectra).  DERIV multiplies by i*kx (zero on the Nyquist column,
// src/field.cpp:310-318).  The imaginary parts of the kx = 0 and kx = NX/2
// columns are dropped, i.e. the line is the Hermitian part, which is what
// the reference keeps when it reads only the real part after an inverse
// complex transform.
template <typename T, int N, bool DERIV>
__device__ __forceinline__ void load_c2r(cplx<T> *v, const cplx<T> *__restrict__ in, size_t rowA, size_t rowB,
                                         int t, const double *__restrict__ kxv)
{
    using S = LineShape<N>;
    using C = cplx<T>;
    if (t < S::N1) {
#pragma unroll
        for (int m = 0; m < S::N2; m++) {
            const int k = t + S::N1 * m;
            const int kk = k <= N / 2 ? k : N - k;
            C a = in[rowA + kk], b = in[rowB + kk];
            if (kk == 0 || kk == N / 2) { a.y = 0; b.y = 0; }
            if constexpr (DERIV) {
                const T kv = kk == N / 2 ? T(0) : (T)kxv[kk];
                a = C{-kv * a.y, kv * a.x};
                b = C{-kv * b.y, kv * b.x};
            }
            if (k <= N / 2) v[m] = C{a.x - b.y, a.y + b.x};      // A + iB
            else v[m] = C{a.x + b.y, -a.y + b.x};                // conj(A) + i conj(B)
        }
    }
}

// After a forward transform of the packed line, thread t < N1 holds
// Z[t + N1*m] (split (N2, N1)).  Write Z to the line buffer in natural order,
// then split into the two half spectra and store rows rowA, rowB.
template <typename T, int N, int HOLD>
__device__ __forceinline__ void store_r2c(cplx<T> *v, cplx<T> *L, cplx<T> *__restrict__ out, size_t rowA,
                                          size_t rowB, int t)
{
    using S = LineShape<N>;
    using C = cplx<T>;
    if (t < N / HOLD) {
#pragma unroll
        for (int m = 0; m < HOLD; m++) L[t + (N / HOLD) * m] = v[m];
    }
    __syncthreads();
    for (int k = t; k <= N / 2; k += S::P) {
        const C zk = L[k], zm = L[(N - k) % N];
        const C A{(zk.x + zm.x) * T(0.5), (zk.y - zm.y) * T(0.5)};
        const C B{(zk.y + zm.y) * T(0.5), (zm.x - zk.x) * T(0.5)};
        out[rowA + k] = A;
        out[rowB + k] = B;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// P1: conv = u . grad c in real space (src/main.cpp:91, field.cpp:392-400),
// then forward x transform.  Also used for the initial c / f transforms.
// ---------------------------------------------------------------------------
template <typename T>
struct ProdConv {
    const T *u0, *u1, *u2, *g0, *g1, *g2;
    __device__ __forceinline__ T operator()(int, size_t i) const
    {
        T a = T(0);
        a += u0[i] * g0[i];
        a += u1[i] * g1[i];
        a += u2[i] * g2[i];
        return a;
    }
};

// one real field as is (stand-alone transforms)
template <typename T>
struct ProdCopy {
    const T *a;
    __device__ __forceinline__ T operator()(int, size_t i) const { return a[i]; }
};

// field 0: c; field 1: f = (lambda/eps^2)(c^3 - c) (src/main.cpp:81)
template <typename T>
struct ProdCF {
    const T *c;
    T fscale;
    __device__ __forceinline__ T operator()(int f, size_t i) const
    {
        const T v = c[i];
        if (f == 0) return v;
        return (v * v * v - v) * fscale;
    }
};

// Fixed-order FP64 sum of one value per thread over the block; the result
// is valid in thread 0.  `red` holds blockDim.x doubles.
__device__ __forceinline__ double block_sum_f64(double v, double *red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    return red[0];
}

// Fixed-order FP64 sums: out[q] = sum_i part[i*stride + q], q < nq (one block).
__global__ __launch_bounds__(256) void k_sum_strided(const double *part, int nb, int stride, int nq, double *out)
{
    __shared__ double s[256];
    for (int q = 0; q < nq; q++) {
        double a = 0;
        for (int i = threadIdx.x; i < nb; i += 256) a += part[(size_t)i * stride + q];
        s[threadIdx.x] = a;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = s[0];
        __syncthreads();
    }
}

// part (optional): FP64 sum of field 0's real values per workgroup, for
// the k = 0 mode of the FP32-mixed path.
template <typename T, int N, int NF, class Prod>
__global__ __launch_bounds__(TILE *LineShape<N>::P) void k_x_forward(Prod prod, cplx<T> *__restrict__ out0,
                                                                      cplx<T> *__restrict__ out1, Dims d,
                                                                      const cplx<T> *__restrict__ tw,
                                                                      double *__restrict__ part)
{
    using S = LineShape<N>;
    using C = cplx<T>;
    __shared__ C lds[TILE * S::LDS];
    const int lp = threadIdx.x / S::P, t = threadIdx.x % S::P;
    const int rl = (blockIdx.x * TILE + lp) * 2;
    C *L = lds + lp * S::LDS;
#pragma unroll
    for (int f = 0; f < NF; f++) {
        C v[S::M];
        double acc = 0.0;
        if (t < S::N1) {
#pragma unroll
            for (int m = 0; m < S::N2; m++) {
                const int x = t + S::N1 * m;
                v[m] = C{prod(f, (size_t)rl * d.nx + x), prod(f, (size_t)(rl + 1) * d.nx + x)};
                acc += (double)v[m].x + (double)v[m].y;
            }
        }
        if (f == 0 && part != nullptr) {
            const double s = block_sum_f64(acc, reinterpret_cast<double *>(lds));
            if (threadIdx.x == 0) part[blockIdx.x] = s;
            __syncthreads();
        }
        fft_line<T, N, S::N1, -1>(v, L, t, tw);
        store_r2c<T, N, S::N1>(v, L, f == 0 ? out0 : out1, (size_t)rl * d.nxp, (size_t)(rl + 1) * d.nxp, t);
    }
}

// ---------------------------------------------------------------------------
// Plain inverse x transform (C2R) of NF fields into real arrays.
// ---------------------------------------------------------------------------
template <typename T, int N, int NF>
__global__ __launch_bounds__(TILE *LineShape<N>::P) void k_x_inverse(const cplx<T> *__restrict__ in0,
                                                                      const cplx<T> *__restrict__ in1,
                                                                      const cplx<T> *__restrict__ in2,
                                                                      T *__restrict__ o0, T *__restrict__ o1,
                                                                      T *__restrict__ o2, Dims d,
                                                                      const cplx<T> *__restrict__ tw)
{
    using S = LineShape<N>;
    using C = cplx<T>;
    __shared__ C lds[TILE * S::LDS];
    const int lp = threadIdx.x / S::P, t = threadIdx.x % S::P;
    const int rl = (blockIdx.x * TILE + lp) * 2;
    C *L = lds + lp * S::LDS;
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const C *in = f == 0 ? in0 : (f == 1 ? in1 : in2);
        T *o = f == 0 ? o0 : (f == 1 ? o1 : o2);
        C v[S::M];
        load_c2r<T, N, false>(v, in, (size_t)rl * d.nxp, (size_t)(rl + 1) * d.nxp, t, nullptr);
        fft_line<T, N, S::N1, +1>(v, L, t, tw);
        if (t < S::N2) {
#pragma unroll
            for (int m = 0; m < S::N1; m++) {
                const int x = t + S::N2 * m;
                o[(size_t)rl * d.nx + x] = v[m].x;
                o[(size_t)(rl + 1) * d.nx + x] = v[m].y;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// P5: inverse x of (c_hat, dy c_hat, dz c_hat) -> c, grad c (real), then
// f = (lambda/eps^2)(c^3 - c) and its forward x transform (src/main.cpp:153).
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ __launch_bounds__(TILE *LineShape<N>::P) void k_x_cgrad_f(
    const cplx<T> *__restrict__ ch, const cplx<T> *__restrict__ dyc, const cplx<T> *__restrict__ dzc,
    T *__restrict__ c, T *__restrict__ gx, T *__restrict__ gy, T *__restrict__ gz, cplx<T> *__restrict__ fout,
    T fscale, Dims d, const double *__restrict__ kxv, const cplx<T> *__restrict__ tw)
{
    using S = LineShape<N>;
    using C = cplx<T>;
    __shared__ C lds[TILE * S::LDS];
    const int lp = threadIdx.x / S::P, t = threadIdx.x % S::P;
    const int rl = (blockIdx.x * TILE + lp) * 2;
    const size_t rA = (size_t)rl * d.nxp, rB = rA + d.nxp;
    const size_t xA = (size_t)rl * d.nx, xB = xA + d.nx;
    C *L = lds + lp * S::LDS;
    C cv[S::M];
    {
        load_c2r<T, N, false>(cv, ch, rA, rB, t, kxv);
        fft_line<T, N, S::N1, +1>(cv, L, t, tw);
        if (t < S::N2) {
#pragma unroll
            for (int m = 0; m < S::N1; m++) {
                const int x = t + S::N2 * m;
                c[xA + x] = cv[m].x;
                c[xB + x] = cv[m].y;
            }
        }
    }
#pragma unroll
    for (int f = 0; f < 3; f++) {
        C v[S::M];
        if (f == 0) load_c2r<T, N, true>(v, ch, rA, rB, t, kxv);
        else load_c2r<T, N, false>(v, f == 1 ? dyc : dzc, rA, rB, t, kxv);
        fft_line<T, N, S::N1, +1>(v, L, t, tw);
        T *g = f == 0 ? gx : (f == 1 ? gy : gz);
        if (t < S::N2) {
#pragma unroll
            for (int m = 0; m < S::N1; m++) {
                const int x = t + S::N2 * m;
                g[xA + x] = v[m].x;
                g[xB + x] = v[m].y;
            }
        }
    }
    // f from c in registers: thread t < N2 holds x = t + N2*m -> forward split (N2, N1)
    if (t < S::N2) {
#pragma unroll
        for (int m = 0; m < S::N1; m++) {
            const T a = cv[m].x, b = cv[m].y;
            cv[m] = C{(a * a * a - a) * fscale, (b * b * b - b) * fscale};
        }
    }
    fft_line<T, N, S::N2, -1>(cv, L, t, tw);
    store_r2c<T, N, S::N2>(cv, L, fout, rA, rB, t);
}

// ---------------------------------------------------------------------------
// P9: inverse x of mu_hat -> mu (real); force F = mu grad c (+ particle
// spread) + buoyancy (src/main.cpp:149-172); forward x transform of F.
// Spread entries for real line l are spr_start[l]..spr_start[l+1], sorted by
// (x, particle id); they are summed per node in that order, then added to
// mu*grad c, then buoyancy is added, which is the reference's association
// ((mu grad c + spread) + buoyancy).
// ---------------------------------------------------------------------------
template <typename T>
struct SpreadEntry {
    int x;
    int pid;
    T v[3];
};

// Spread sum of one grid node: x on its line and the three force components.
template <typename T>
struct SpreadNode {
    int x;
    T v[3];
};

// P9 (see the pass list above).  nodes[line_start[l] + k], k < node_cnt[l],
// are the per-node spread sums of real line l (unique x per line), each
// already summed in ascending particle order; every thread adds whole
// nodes, so there is one add per node and no ordering between threads.
// Buoyancy enters the z component only: g = (0, 0, gravity)
// (field.cpp:122-133).
template <typename T, int N>
__global__ __launch_bounds__(TILE *LineShape<N>::P) void k_x_force(
    cplx<T> *__restrict__ mu_io /* in: mu_hat (X); out: F_0 (X) */, cplx<T> *__restrict__ f1,
    cplx<T> *__restrict__ f2, const T *__restrict__ c, const T *__restrict__ gx, const T *__restrict__ gy,
    const T *__restrict__ gz, const int *__restrict__ line_start, const int *__restrict__ node_cnt,
    const SpreadNode<T> *__restrict__ nodes, int use_spread, T rho1, T rho2, T gravity, Dims d,
    const cplx<T> *__restrict__ tw)
{
    using S = LineShape<N>;
    using C = cplx<T>;
    __shared__ C lds[TILE * S::LDS];
    const int lp = threadIdx.x / S::P, t = threadIdx.x % S::P;
    const int rl = (blockIdx.x * TILE + lp) * 2;
    const size_t rA = (size_t)rl * d.nxp, rB = rA + d.nxp;
    const size_t xA = (size_t)rl * d.nx, xB = xA + d.nx;
    C *L = lds + lp * S::LDS;
    C muv[S::M];
    load_c2r<T, N, false>(muv, mu_io, rA, rB, t, nullptr);
    fft_line<T, N, S::N1, +1>(muv, L, t, tw);
    // thread t < N2 holds mu at x = t + N2*m for lines rl (x part) and rl+1 (y part)
#pragma unroll
    for (int f = 0; f < 3; f++) {
        const T *g = f == 0 ? gx : (f == 1 ? gy : gz);
        C v[S::M];
        if (t < S::N2) {
#pragma unroll
            for (int m = 0; m < S::N1; m++) {
                const int x = t + S::N2 * m;
                v[m] = C{muv[m].x * g[xA + x], muv[m].y * g[xB + x]};
            }
        }
        if (use_spread) {
            T *Lr = reinterpret_cast<T *>(L);  // natural-order buffer: line rl then rl+1
            if (t < S::N2) {
#pragma unroll
                for (int m = 0; m < S::N1; m++) {
                    const int x = t + S::N2 * m;
                    Lr[x] = v[m].x;
                    Lr[N + x] = v[m].y;
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int b = line_start[rl + q], cnt = node_cnt[rl + q];
                for (int k = t; k < cnt; k += S::P) {
                    const SpreadNode<T> nd = nodes[b + k];
                    Lr[q * N + nd.x] += nd.v[f];
                }
            }
            __syncthreads();
            if (t < S::N2) {
#pragma unroll
                for (int m = 0; m < S::N1; m++) {
                    const int x = t + S::N2 * m;
                    v[m] = C{Lr[x], Lr[N + x]};
                }
            }
            __syncthreads();
        }
        if (f == 2 && t < S::N2) {
            // + ((1+c)/2*rho1 + (1-c)/2*rho2) * gravity, src/main.cpp:172
#pragma unroll
            for (int m = 0; m < S::N1; m++) {
                const int x = t + S::N2 * m;
                const T ca = c[xA + x], cb = c[xB + x];
                const T ba = (ca + T(1)) / T(2) * rho1 + (ca * T(-1) + T(1)) / T(2) * rho2;
                const T bb = (cb + T(1)) / T(2) * rho1 + (cb * T(-1) + T(1)) / T(2) * rho2;
                v[m] = C{v[m].x + ba * gravity, v[m].y + bb * gravity};
            }
        }
        fft_line<T, N, S::N2, -1>(v, L, t, tw);
        store_r2c<T, N, S::N2>(v, L, f == 0 ? mu_io : (f == 1 ? f1 : f2), rA, rB, t);
    }
}

// ---------------------------------------------------------------------------
// y pass: transforms along y of NF fields in place.  GRAD (inverse only):
// field 0 also yields i*ky*field0 into `extra` (zero on ky = NY/2).
// ---------------------------------------------------------------------------
template <typename T, int N, int SIGN, int NF, bool GRAD>
__global__ __launch_bounds__(TILE *LineShape<N>::P) void k_y(cplx<T> *__restrict__ a0, cplx<T> *__restrict__ a1,
                                                             cplx<T> *__restrict__ a2, cplx<T> *__restrict__ extra,
                                                             Dims d, const double *__restrict__ kyv,
                                                             const cplx<T> *__restrict__ tw)
{
    using S = LineShape<N>;
    using C = cplx<T>;
    __shared__ C lds[TILE * S::LDS];
    const int lane = threadIdx.x % TILE, t = threadIdx.x / TILE;
    const int kx = blockIdx.x * TILE + lane, z = blockIdx.y;
    const bool ok = kx < d.nxh;
    C *L = lds + lane * S::LDS;
    const size_t base = hidx(d, ok ? kx : 0, 0, z);
    const size_t st = d.nxp;
#pragma unroll
    for (int f = 0; f < NF; f++) {
        C *a = f == 0 ? a0 : (f == 1 ? a1 : a2);
        C v[S::M];
        if (t < S::N1) {
#pragma unroll
            for (int m = 0; m < S::N2; m++) v[m] = ok ? a[base + st * (t + S::N1 * m)] : C{0, 0};
        }
        if (GRAD && f == 0) {
            C w[S::M];
            if (t < S::N1) {
#pragma unroll
                for (int m = 0; m < S::N2; m++) {
                    const int y = t + S::N1 * m;
                    const T kv = y == N / 2 ? T(0) : (T)kyv[y];
                    w[m] = C{-kv * v[m].y, kv * v[m].x};
                }
            }
            fft_line<T, N, S::N1, SIGN>(w, L, t, tw);
            if (ok && t < S::N2) {
#pragma unroll
                for (int m = 0; m < S::N1; m++) extra[base + st * (t + S::N2 * m)] = w[m];
            }
        }
        fft_line<T, N, S::N1, SIGN>(v, L, t, tw);
        if (ok && t < S::N2) {
#pragma unroll
            for (int m = 0; m < S::N1; m++) a[base + st * (t + S::N2 * m)] = v[m];
        }
    }
}

// ---------------------------------------------------------------------------
// z hubs.  Each loads NIN XY lines (z direction) in D(N1), forward
// transforms (result in D(N2): thread t holds kz = t + N2*m), applies the
// k-space operation on registers, inverse transforms NOUT lines and stores
// them in D(N1) in place.
// ---------------------------------------------------------------------------
struct KParams {
    double dt, mobility, lambda, mu;
    double inv_g;   // 1/(nx*ny*nz), the inverse transform's normalisation
    int first;      // 1 on the first step (first-order start, main.cpp:95-114)
    int use_conv;   // switch_ns
};

template <typename T, int N>
__device__ __forceinline__ void hub_load(cplx<T> *v, const cplx<T> *__restrict__ a, size_t base, size_t st, int t,
                                         bool ok)
{
    using S = LineShape<N>;
    if (t < S::N1) {
#pragma unroll
        for (int m = 0; m < S::N2; m++) v[m] = ok ? a[base + st * (t + S::N1 * m)] : cplx<T>{0, 0};
    }
}

template <typename T, int N>
__device__ __forceinline__ void hub_store(const cplx<T> *v, cplx<T> *__restrict__ a, size_t base, size_t st, int t,
                                          bool ok)
{
    using S = LineShape<N>;
    if (ok && t < S::N1) {
#pragma unroll
        for (int m = 0; m < S::N2; m++) a[base + st * (t + S::N1 * m)] = v[m];
    }
}

// P3: CH update.  conv (XY) -> conv_hat; with c_hat^n, f_hat^n and the
// history term hist = -c^{n-1} + 2dtM k^2 f^{n-1} + 2dt conv^{n-1}:
//   c^{n+1} = (4c^n - 4dtM k^2 f^n - 4dt conv^n + hist) / (3 + 2dtM lambda k^4)
// (src/main.cpp:117-127) or on the first step
//   c^1 = (c^0 - dtM k^2 f^0 - dt conv^0) / (1 + dtM lambda k^4)  (:106-113).
// Outputs c^{n+1} (K), the new history (K), and c^{n+1}/G, i kz c^{n+1}/G
// inverse-z transformed (XY) for the gradient and real-space c.
// FP64 k = 0 mode (FP32-mixed path): k0[cur] = c_hat^n(0), k0[cur^1] gets
// c_hat^{n+1}(0), k0[2] = history(0); conv0 = FP64 sum of conv (its k = 0
// transform).  k0 == nullptr: the k = 0 mode is carried like every other.
struct K0 {
    double *k0;
    const double *conv0;
    int cur;
};

template <typename T, int N>
__global__ __launch_bounds__(TILE *LineShape<N>::P) void k_hub_ch(
    cplx<T> *__restrict__ conv_io /* in conv XY, out c XY */, cplx<T> *__restrict__ dzc_out,
    const cplx<T> *__restrict__ c_now, cplx<T> *__restrict__ c_next, const cplx<T> *__restrict__ f_now,
    cplx<T> *__restrict__ hist, KParams kp, Dims d, const double *__restrict__ kxv, const double *__restrict__ kyv,
    const double *__restrict__ kzv, const cplx<T> *__restrict__ tw, K0 k0)
{
    using S = LineShape<N>;
    using C = cplx<T>;
    __shared__ C lds[TILE * S::LDS];
    const int lane = threadIdx.x % TILE, t = threadIdx.x / TILE;
    const int kx = blockIdx.x * TILE + lane, y = blockIdx.y;
    const bool ok = kx < d.nxh;
    C *L = lds + lane * S::LDS;
    const size_t base = hidx(d, ok ? kx : 0, y, 0);
    const size_t st = (size_t)d.nxp * d.ny;
    C v[S::M];
    if (kp.use_conv) {
        hub_load<T, N>(v, conv_io, base, st, t, ok);
        fft_line<T, N, S::N1, -1>(v, L, t, tw);
    } else if (t < S::N2) {
#pragma unroll
        for (int m = 0; m < S::N1; m++) v[m] = C{0, 0};
    }
    C w[S::M];
    if (t < S::N2) {
        const double kxx = kxv[ok ? kx : 0], kyy = kyv[y];
#pragma unroll
        for (int m = 0; m < S::N1; m++) {
            const int z = t + S::N2 * m;
            const double kzz = kzv[z];
            double k2 = 0.0;
            k2 += kxx * kxx;
            k2 += kyy * kyy;
            k2 += kzz * kzz;
            const size_t i = base + st * z;
            C cn = C{0, 0}, fn = C{0, 0}, h = C{0, 0};
            if (ok) { cn = c_now[i]; fn = f_now[i]; h = hist[i]; }
            const C cv = v[m];
            C nc;
            // The update is evaluated in FP64 from the stored values (FP32 in
            // the mixed path): coefficients rounded to FP32 would bias every
            // mode's growth factor by ~1e-7 per step, and the bias compounds
            // over the run; rounding only the result is unbiased.
            const double cnx = cn.x, cny = cn.y, fnx = fn.x, fny = fn.y, cvx = cv.x, cvy = cv.y;
            if (kp.first) {
                const double m1 = k2 * (kp.dt * kp.mobility);
                const double div = k2 * k2 * (kp.dt * kp.mobility * kp.lambda) + 1.0;
                double r = cnx - m1 * fnx, im = cny - m1 * fny;
                r -= cvx * kp.dt;
                im -= cvy * kp.dt;
                nc = C{(T)(r / div), (T)(im / div)};
            } else {
                const double m4 = k2 * (4.0 * kp.dt * kp.mobility);
                const double b4 = 4.0 * kp.dt;
                const double div = k2 * k2 * (2.0 * kp.dt * kp.mobility * kp.lambda) + 3.0;
                double r = cnx * 4.0 - m4 * fnx, im = cny * 4.0 - m4 * fny;
                r -= cvx * b4;
                im -= cvy * b4;
                r += (double)h.x;
                im += (double)h.y;
                nc = C{(T)(r / div), (T)(im / div)};
            }
            const double m2 = k2 * (2.0 * kp.dt * kp.mobility);
            const double b2 = 2.0 * kp.dt;
            C nh = C{(T)(-cnx + m2 * fnx + cvx * b2), (T)(-cny + m2 * fny + cvy * b2)};
            const T ig = (T)kp.inv_g;
            if (k0.k0 != nullptr && ok && kx == 0 && y == 0 && z == 0) {
                // mass mode in FP64: k^2 = 0, so only c, conv and history enter
                const double c0 = k0.k0[k0.cur], h0 = k0.k0[2];
                const double cv0 = kp.use_conv ? *k0.conv0 : 0.0;
                double n0;
                if (kp.first) n0 = c0 - cv0 * kp.dt;
                else n0 = (c0 * 4.0 - cv0 * (4.0 * kp.dt) + h0) / 3.0;
                k0.k0[k0.cur ^ 1] = n0;
                k0.k0[2] = -c0 + cv0 * (2.0 * kp.dt);
                nc = C{(T)n0, T(0)};
                nh = C{(T)k0.k0[2], T(0)};
                v[m] = C{(T)(n0 * kp.inv_g), T(0)};
            } else {
                v[m] = C{nc.x * ig, nc.y * ig};
            }
            if (ok) { c_next[i] = nc; hist[i] = nh; }
            const T kz = z == N / 2 ? T(0) : (T)kzz;
            w[m] = C{-kz * v[m].y, kz * v[m].x};
        }
    }
    fft_line<T, N, S::N2, +1>(v, L, t, tw);
    hub_store<T, N>(v, conv_io, base, st, t, ok);
    fft_line<T, N, S::N2, +1>(w, L, t, tw);
    hub_store<T, N>(w, dzc_out, base, st, t, ok);
}

// Initial hub: (c XY, f XY) -> c_hat^0, f_hat^0 (K) and c/G, i kz c/G (XY).
template <typename T, int N>
__global__ __launch_bounds__(TILE *LineShape<N>::P) void k_hub_init(cplx<T> *__restrict__ c_io,
                                                                    cplx<T> *__restrict__ f_io /* out: dz c XY */,
                                                                    cplx<T> *__restrict__ c_k,
                                                                    cplx<T> *__restrict__ f_k, KParams kp, Dims d,
                                                                    const double *__restrict__ kzv,
                                                                    const cplx<T> *__restrict__ tw)
{
    using S = LineShape<N>;
    using C = cplx<T>;
    __shared__ C lds[TILE * S::LDS];
    const int lane = threadIdx.x % TILE, t = threadIdx.x / TILE;
    const int kx = blockIdx.x * TILE + lane, y = blockIdx.y;
    const bool ok = kx < d.nxh;
    C *L = lds + lane * S::LDS;
    const size_t base = hidx(d, ok ? kx : 0, y, 0);
    const size_t st = (size_t)d.nxp * d.ny;
    C fv[S::M];
    hub_load<T, N>(fv, f_io, base, st, t, ok);
    fft_line<T, N, S::N1, -1>(fv, L, t, tw);
    if (ok && t < S::N2) {
#pragma unroll
        for (int m = 0; m < S::N1; m++) f_k[base + st * (t + S::N2 * m)] = fv[m];
    }
    C v[S::M], w[S::M];
    hub_load<T, N>(v, c_io, base, st, t, ok);
    fft_line<T, N, S::N1, -1>(v, L, t, tw);
    if (t < S::N2) {
        const T ig = (T)kp.inv_g;
#pragma unroll
        for (int m = 0; m < S::N1; m++) {
            const int z = t + S::N2 * m;
            if (ok) c_k[base + st * z] = v[m];
            v[m] = C{v[m].x * ig, v[m].y * ig};
            const T kz = z == N / 2 ? T(0) : (T)kzv[z];
            w[m] = C{-kz * v[m].y, kz * v[m].x};
        }
    }
    fft_line<T, N, S::N2, +1>(v, L, t, tw);
    hub_store<T, N>(v, c_io, base, st, t, ok);
    fft_line<T, N, S::N2, +1>(w, L, t, tw);
    hub_store<T, N>(w, f_io, base, st, t, ok);
}

// P7: f (XY) -> f_hat^{n+1} (K); mu_hat = lambda k^2 c_hat + f_hat
// (src/main.cpp:157) -> mu/G (XY, in place of f).
template <typename T, int N>
__global__ __launch_bounds__(TILE *LineShape<N>::P) void k_hub_mu(cplx<T> *__restrict__ f_io,
                                                                  cplx<T> *__restrict__ f_k,
                                                                  const cplx<T> *__restrict__ c_k, KParams kp,
                                                                  Dims d, const double *__restrict__ kxv,
                                                                  const double *__restrict__ kyv,
                                                                  const double *__restrict__ kzv,
                                                                  const cplx<T> *__restrict__ tw)
{
    using S = LineShape<N>;
    using C = cplx<T>;
    __shared__ C lds[TILE * S::LDS];
    const int lane = threadIdx.x % TILE, t = threadIdx.x / TILE;
    const int kx = blockIdx.x * TILE + lane, y = blockIdx.y;
    const bool ok = kx < d.nxh;
    C *L = lds + lane * S::LDS;
    const size_t base = hidx(d, ok ? kx : 0, y, 0);
    const size_t st = (size_t)d.nxp * d.ny;
    C v[S::M];
    hub_load<T, N>(v, f_io, base, st, t, ok);
    fft_line<T, N, S::N1, -1>(v, L, t, tw);
    if (t < S::N2) {
        const double kxx = kxv[ok ? kx : 0], kyy = kyv[y];
#pragma unroll
        for (int m = 0; m < S::N1; m++) {
            const int z = t + S::N2 * m;
            const double kzz = kzv[z];
            double k2 = 0.0;
            k2 += kxx * kxx;
            k2 += kyy * kyy;
            k2 += kzz * kzz;
            const size_t i = base + st * z;
            const C fh = v[m];
            C cc = C{0, 0};
            if (ok) { f_k[i] = fh; cc = c_k[i]; }
            // FP64 arithmetic on the stored values, one rounding at the end
            const double ml = k2 * kp.lambda;
            v[m] = C{(T)((ml * (double)cc.x + (double)fh.x) * kp.inv_g),
                     (T)((ml * (double)cc.y + (double)fh.y) * kp.inv_g)};
        }
    }
    fft_line<T, N, S::N2, +1>(v, L, t, tw);
    hub_store<T, N>(v, f_io, base, st, t, ok);
}

// P11: Stokes, u_hat = (F - k (k.F)/k^2)/(mu k^2), 0 at k = 0
// (src/main.cpp:177-178, division by zero gives zero, field.cpp:1010-1019).
// The reference keeps the real part of the inverse complex transform; its
// projection is not Hermitian where a wave-number component sits on a
// Nyquist index (k_d(N/2) = +pi/dx on both sides).  The real part equals the
// inverse of the Hermitian part, which drops the k_d k_e cross terms
// between Nyquist and non-Nyquist components; that is what is formed here.
template <typename T, int N>
__global__ __launch_bounds__(TILE *LineShape<N>::P) void k_hub_stokes(cplx<T> *__restrict__ a0,
                                                                      cplx<T> *__restrict__ a1,
                                                                      cplx<T> *__restrict__ a2, KParams kp, Dims d,
                                                                      const double *__restrict__ kxv,
                                                                      const double *__restrict__ kyv,
                                                                      const double *__restrict__ kzv,
                                                                      const cplx<T> *__restrict__ tw)
{
    using S = LineShape<N>;
    using C = cplx<T>;
    __shared__ C lds[TILE * S::LDS];
    const int lane = threadIdx.x % TILE, t = threadIdx.x / TILE;
    const int kx = blockIdx.x * TILE + lane, y = blockIdx.y;
    const bool ok = kx < d.nxh;
    C *L = lds + lane * S::LDS;
    const size_t base = hidx(d, ok ? kx : 0, y, 0);
    const size_t st = (size_t)d.nxp * d.ny;
    C v0[S::M], v1[S::M], v2[S::M];
    // all three loads in flight before the first transform
    hub_load<T, N>(v0, a0, base, st, t, ok);
    hub_load<T, N>(v1, a1, base, st, t, ok);
    hub_load<T, N>(v2, a2, base, st, t, ok);
    fft_line<T, N, S::N1, -1>(v0, L, t, tw);
    fft_line<T, N, S::N1, -1>(v1, L, t, tw);
    fft_line<T, N, S::N1, -1>(v2, L, t, tw);
    if (t < S::N2) {
        const int kxi = ok ? kx : 0;
        const double kxx = kxv[kxi], kyy = kyv[y];
        const bool nx_ = kxi == d.nx / 2, ny_ = y == d.ny / 2;
#pragma unroll
        for (int m = 0; m < S::N1; m++) {
            const int z = t + S::N2 * m;
            const double kzz = kzv[z];
            const bool nz_ = z == N / 2;
            double k2 = 0.0;
            k2 += kxx * kxx;
            k2 += kyy * kyy;
            k2 += kzz * kzz;
            const double kv[3] = {kxx, kyy, kzz};
            const bool nq[3] = {nx_, ny_, nz_};
            const double F[3][2] = {{v0[m].x, v0[m].y}, {v1[m].x, v1[m].y}, {v2[m].x, v2[m].y}};
            // k.F split by Nyquist class (FP64 arithmetic on the stored values)
            double sar = 0, sai = 0, sbr = 0, sbi = 0;
#pragma unroll
            for (int e = 0; e < 3; e++) {
                if (nq[e]) { sbr += kv[e] * F[e][0]; sbi += kv[e] * F[e][1]; }
                else { sar += kv[e] * F[e][0]; sai += kv[e] * F[e][1]; }
            }
            const double vis = k2 * kp.mu;
            C U[3];
#pragma unroll
            for (int e = 0; e < 3; e++) {
                const double sr = nq[e] ? sbr : sar, si = nq[e] ? sbi : sai;
                double pr = kv[e] * sr, pi_ = kv[e] * si;
                if (k2 != 0.0) { pr /= k2; pi_ /= k2; } else { pr = 0.0; pi_ = 0.0; }
                double ur = F[e][0] - pr, ui = F[e][1] - pi_;
                if (vis != 0.0) { ur /= vis; ui /= vis; } else { ur = 0.0; ui = 0.0; }
                U[e] = C{(T)(ur * kp.inv_g), (T)(ui * kp.inv_g)};
            }
            v0[m] = U[0];
            v1[m] = U[1];
            v2[m] = U[2];
        }
    }
    fft_line<T, N, S::N2, +1>(v0, L, t, tw);
    hub_store<T, N>(v0, a0, base, st, t, ok);
    fft_line<T, N, S::N2, +1>(v1, L, t, tw);
    hub_store<T, N>(v1, a1, base, st, t, ok);
    fft_line<T, N, S::N2, +1>(v2, L, t, tw);
    hub_store<T, N>(v2, a2, base, st, t, ok);
}

// Plain forward / inverse z transform of one field in place (unit tests).
template <typename T, int N, int SIGN>
__global__ __launch_bounds__(TILE *LineShape<N>::P) void k_z(cplx<T> *__restrict__ a, Dims d,
                                                             const cplx<T> *__restrict__ tw)
{
    using S = LineShape<N>;
    using C = cplx<T>;
    __shared__ C lds[TILE * S::LDS];
    const int lane = threadIdx.x % TILE, t = threadIdx.x / TILE;
    const int kx = blockIdx.x * TILE + lane, y = blockIdx.y;
    const bool ok = kx < d.nxh;
    C *L = lds + lane * S::LDS;
    const size_t base = hidx(d, ok ? kx : 0, y, 0);
    const size_t st = (size_t)d.nxp * d.ny;
    C v[S::M];
    hub_load<T, N>(v, a, base, st, t, ok);
    fft_line<T, N, S::N1, SIGN>(v, L, t, tw);
    // thread t < N2 holds z = t + N2*m
    if (ok && t < S::N2) {
#pragma unroll
        for (int m = 0; m < S::N1; m++) a[base + st * (t + S::N2 * m)] = v[m];
    }
}

}  // namespace fipi
