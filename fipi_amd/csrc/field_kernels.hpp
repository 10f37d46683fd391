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
//            the reverse (C2R).  128 (FP64) or 256 (FP32) threads per
//            workgroup.
//   y pass   a 64- or 128-byte tile of neighbouring kx columns x all y at
//            one z.
//   z hub    a 64- or 128-byte tile of kx columns x all z at one y: forward
//            z transform, the fused k-space update, inverse z transform,
//            all in registers.
// Every line is held in the distribution D of fft_core.hpp (thread t of
// the line holds elements t + q*P), before and after each transform.
// Wave numbers are read from per-axis tables computed on the host with the
// reference expression i*2.0*pi/L (src/field.cpp:108-115).
#pragma once
#include "fft_core.hpp"

namespace fipi {

// nxh = nx/2 + 1 spectral columns; nxp >= nxh is the row pitch of every
// half-spectrum array, a multiple of 128 bytes so that the 16-column tiles
// of the y and z passes start on cache-line boundaries.
// A context holds a slab of equal boxes: box b's arrays start b*rs (real)
// and b*hs (half spectrum) elements after box 0's; every field kernel takes
// its box from blockIdx.z.
struct Dims {
    int nx, ny, nz, nxh, nxp;
    size_t rs, hs;
};

// threads per x-pass workgroup (at most): FP64 kernels hold ~230 VGPRs,
// so a CU runs 8 of their waves whatever the workgroup size; 2-wave
// workgroups let 4 of them work in different phases (load / FFT / store)
template <typename T> struct XT_of {
    static constexpr int v = sizeof(T) == 8 ? 128 : 256;
};

// x-pass workgroup: up to LPW complex lines (2 real lines each) of P
// threads; a launch uses blockDim.x / P lines, a divisor of ny*nz/2, so
// every line of every workgroup exists
template <typename T, int N, int EV = PlanE_for<T, N>::E> struct XShape {
    static constexpr int LPW = XT_of<T>::v / Line<T, N, EV>::P;
};

// Values per thread of the velocity/convection x pass: FP64 256-point rows
// use the 4.4.4.4 plan there (measured faster than 8.4.8 for that pass and
// slower for the force pass and the Stokes hub, which keep the default).
// FP32 keeps the default: its per-workgroup partial sums of conv (the
// FP64 k = 0 mode) are laid out by the default x-pass grid.
template <typename T, int N> struct XE_conv {
    static constexpr int v = (sizeof(T) == 8 && N == 256) ? 4 : PlanE_for<T, N>::E;
};
static_assert(XE_conv<float, 256>::v == PlanE_for<float, 256>::E, "FP32 conv partials use the default grid");

// kx columns per y/z-pass workgroup: a TB-byte run of complex values
// (TB = 64 or 128, chosen per launch site and edge length)
template <typename T, int TB> struct TileW {
    static constexpr int L = TB / (int)sizeof(cplx<T>);
};

__device__ __forceinline__ size_t hidx(const Dims &d, int kx, int y, int z)
{
    return (size_t)kx + (size_t)d.nxp * ((size_t)y + (size_t)d.ny * z);
}

// ---------------------------------------------------------------------------
// x pass helpers
// ---------------------------------------------------------------------------

// Load the packed complex line Z = A + iB of real lines (rl, rl+1) in D:
// thread t gets Z[t + q P].  A = in[rowA..], B = in[rowB..] (half spectra).
// DERIV multiplies by i*kx (zero on the Nyquist column, src/field.cpp:310-318).
// The imaginary parts of the kx = 0 and kx = NX/2 columns are dropped, i.e.
// the line is the Hermitian part, which is what the reference keeps when it
// reads only the real part after an inverse complex transform.
template <typename T, int N, bool DERIV, int EV = PlanE_for<T, N>::E>
__device__ __forceinline__ void load_c2r(cplx<T> *v, const cplx<T> *__restrict__ in, size_t rowA, size_t rowB,
                                         int t, const double *__restrict__ kxv)
{
    constexpr int E = EV, P = Line<T, N, EV>::P;
    using C = cplx<T>;
#pragma unroll
    for (int q = 0; q < E; q++) {
        const int k = t + q * P;
        const int kk = k <= N / 2 ? k : N - k;
        C a = in[rowA + kk], b = in[rowB + kk];
        if (kk == 0 || kk == N / 2) { a.y = 0; b.y = 0; }
        if constexpr (DERIV) {
            const T kv = kk == N / 2 ? T(0) : (T)kxv[kk];
            a = C{-kv * a.y, kv * a.x};
            b = C{-kv * b.y, kv * b.x};
        }
        if (k <= N / 2) v[q] = C{a.x - b.y, a.y + b.x};      // A + iB
        else v[q] = C{a.x + b.y, -a.y + b.x};                // conj(A) + i conj(B)
    }
}

// After a forward transform of the packed line (Z in D), write Z to the
// line buffer in natural order, split into the two half spectra and store
// rows rowA, rowB.
template <typename T, int N, int EV = PlanE_for<T, N>::E>
__device__ __forceinline__ void store_r2c(const cplx<T> *v, cplx<T> *L, cplx<T> *__restrict__ out, size_t rowA,
                                          size_t rowB, int t)
{
    constexpr int E = EV, P = Line<T, N, EV>::P;
    using C = cplx<T>;
#pragma unroll
    for (int q = 0; q < E; q++) L[t + q * P] = v[q];
    __syncthreads();
    for (int k = t; k <= N / 2; k += P) {
        const C zk = L[k], zm = L[(N - k) % N];
        const C A{(zk.x + zm.x) * T(0.5), (zk.y - zm.y) * T(0.5)};
        const C B{(zk.y + zm.y) * T(0.5), (zm.x - zk.x) * T(0.5)};
        out[rowA + k] = A;
        out[rowB + k] = B;
    }
    __syncthreads();
}

// common x-pass geometry: complex line lp of the workgroup = real lines
// rl, rl+1 of box blockIdx.z
#define FIPI_XPASS_SETUP_E(T, N, EV)                                                     \
    constexpr int E = Line<T, N, EV>::E, P = Line<T, N, EV>::P, LL = Line<T, N, EV>::LL; \
    constexpr int LPW = XShape<T, N, EV>::LPW;                                           \
    using C = cplx<T>;                                                       \
    __shared__ C lds[LPW * LL];                                              \
    const int lp = threadIdx.x / P, t = threadIdx.x % P;                     \
    const int rl = (blockIdx.x * (blockDim.x / P) + lp) * 2;                 \
    const size_t zb = blockIdx.z;                                            \
    C *L = lds + lp * LL;                                                    \
    (void)E;                                                                 \
    (void)zb;
#define FIPI_XPASS_SETUP(T, N) FIPI_XPASS_SETUP_E(T, N, (PlanE_for<T, N>::E))

// The twiddle table of the N-point lines into LDS, read there by every
// transform of the kernel instead of from global memory (each transform
// otherwise loads its twiddles after an LDS barrier, inside its dependency
// chain; the force pass went from 0.442 to 0.407 ms FP64, 0.269 to 0.228 ms
// FP32, the one-field y passes 4-5 % faster).  The first barrier of the
// first transform orders these writes before any read.  The start-up pass
// k_x_cgrad_f keeps the global table: there the change costs a wave per
// SIMD.
#define FIPI_TW_TO_LDS_IF(T, N, tw, COND)                                         \
    __shared__ cplx<T> twl_[(COND) ? (N) : 1];                                    \
    if constexpr (COND) {                                                         \
        for (int i_ = threadIdx.x; i_ < (N); i_ += blockDim.x) twl_[i_] = tw[i_]; \
        tw = twl_;                                                                \
    }
#define FIPI_TW_TO_LDS(T, N, tw) FIPI_TW_TO_LDS_IF(T, N, tw, true)

// pointer of box zb in a slab with per-box stride st (null stays null)
template <typename P_>
__device__ __forceinline__ P_ *boxp(P_ *p, size_t zb, size_t st)
{
    return p ? p + zb * st : p;
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
// W (optional): wetting source added to f (calibration extension, zero in
// the reference)
struct ProdCF {
    const T *c;
    T fscale;
    const T *W;
    __device__ __forceinline__ T operator()(int f, size_t i) const
    {
        const T v = c[i];
        if (f == 0) return v;
        T r = (v * v * v - v) * fscale;
        if (W) r += W[i];
        return r;
    }
};

// Fixed-order FP64 sum of one value per thread over the block; the result
// is valid in thread 0.  `red` holds blockDim.x rounded up to a power of two
// doubles (at least 2).
__device__ __forceinline__ double block_sum_f64(double v, double *red)
{
    int np = 2;
    while (np < (int)blockDim.x) np <<= 1;
    red[threadIdx.x] = v;
    for (int i = blockDim.x + threadIdx.x; i < np; i += blockDim.x) red[i] = 0.0;
    __syncthreads();
    for (int o = np / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    return red[0];
}

// Fixed-order FP64 sums: out[q] = sum_i part[i*stride + q], q < nq; one
// block per box (box b: part + b*nb*stride, out + b*nq).
__global__ __launch_bounds__(256) void k_sum_strided(const double *part, int nb, int stride, int nq, double *out)
{
    __shared__ double s[256];
    part += (size_t)blockIdx.x * nb * stride;
    out += (size_t)blockIdx.x * nq;
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
__global__ __launch_bounds__(XT_of<T>::v) void k_x_forward(Prod prod, cplx<T> *__restrict__ out0, cplx<T> *__restrict__ out1,
                                                  Dims d, const cplx<T> *__restrict__ tw, double *__restrict__ part)
{
    FIPI_XPASS_SETUP(T, N)
    FIPI_TW_TO_LDS(T, N, tw)
#pragma unroll
    for (int f = 0; f < NF; f++) {
        C v[E];
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int x = t + q * P;
            v[q] = C{prod(f, zb * d.rs + (size_t)rl * d.nx + x), prod(f, zb * d.rs + (size_t)(rl + 1) * d.nx + x)};
            acc += (double)v[q].x + (double)v[q].y;
        }
        if (f == 0 && part != nullptr) {
            const double s = block_sum_f64(acc, reinterpret_cast<double *>(lds));
            if (threadIdx.x == 0) part[zb * gridDim.x + blockIdx.x] = s;
            __syncthreads();
        }
        fft_dist<T, N, -1>(v, L, t, tw);
        store_r2c<T, N>(v, L, boxp(f == 0 ? out0 : out1, zb, d.hs), (size_t)rl * d.nxp, (size_t)(rl + 1) * d.nxp,
                        t);
    }
}

// Plain inverse x transform (C2R) of NF fields into real arrays.
// EV: the plan (values per thread), as the kernel whose output it repeats.
template <typename T, int N, int NF, int EV = PlanE_for<T, N>::E>
__global__ __launch_bounds__(XT_of<T>::v) void k_x_inverse(const cplx<T> *__restrict__ in0, const cplx<T> *__restrict__ in1,
                                                  const cplx<T> *__restrict__ in2, T *__restrict__ o0,
                                                  T *__restrict__ o1, T *__restrict__ o2, Dims d,
                                                  const cplx<T> *__restrict__ tw)
{
    FIPI_XPASS_SETUP_E(T, N, EV)
    FIPI_TW_TO_LDS(T, N, tw)
    const size_t rA = zb * d.hs + (size_t)rl * d.nxp, rB = rA + d.nxp;
    const size_t xA = zb * d.rs + (size_t)rl * d.nx, xB = xA + d.nx;
    C v[NF][E];
#pragma unroll
    for (int f = 0; f < NF; f++) load_c2r<T, N, false, EV>(v[f], f == 0 ? in0 : (f == 1 ? in1 : in2), rA, rB, t, nullptr);
#pragma unroll
    for (int f = 0; f < NF; f++) {
        T *o = f == 0 ? o0 : (f == 1 ? o1 : o2);
        fft_dist<T, N, +1, 1, EV>(v[f], L, t, tw);
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int x = t + q * P;
            o[xA + x] = v[f][q].x;
            o[xB + x] = v[f][q].y;
        }
    }
}

// P13 + next step's P1: inverse x of u_hat (X) -> u (real), then
// conv = u . grad c (src/main.cpp:91, field.cpp:392-400, same sum order as
// ProdConv) and its forward x transform into hout (X): h0, in place of
// u_hat_x, or another array, which leaves u_hat intact.  u0 == nullptr: no
// real u is stored (it lives in registers only).  g holds grad c^{n+1},
// the gradient the next step's convection uses.  part (optional): FP64
// block sums of conv, as k_x_forward.
template <typename T, int N, int EV = XE_conv<T, N>::v>
__global__ __launch_bounds__(XT_of<T>::v) void k_x_inv_conv(const cplx<T> *h0, const cplx<T> *h1, const cplx<T> *h2,
                                                   cplx<T> *hout, T *__restrict__ u0,
                                                   T *__restrict__ u1, T *__restrict__ u2,
                                                   const T *__restrict__ g0, const T *__restrict__ g1,
                                                   const T *__restrict__ g2, Dims d, const cplx<T> *__restrict__ tw,
                                                   double *__restrict__ part)
{
    FIPI_XPASS_SETUP_E(T, N, EV)
    FIPI_TW_TO_LDS(T, N, tw)
    const size_t rA = zb * d.hs + (size_t)rl * d.nxp, rB = rA + d.nxp;
    const size_t xA = zb * d.rs + (size_t)rl * d.nx, xB = xA + d.nx;
    // FP32: all three components in flight from the start; FP64: the next
    // component in flight while one is transformed (three would not fit in
    // the registers of two waves per SIMD)
    constexpr int NV = sizeof(T) == 4 ? 3 : 2;
    C v[NV][E];
#pragma unroll
    for (int f = 0; f < NV; f++) load_c2r<T, N, false, EV>(v[f], f == 0 ? (const C *)h0 : (f == 1 ? h1 : h2), rA, rB, t, nullptr);
    C cv[E];
#pragma unroll
    for (int q = 0; q < E; q++) cv[q] = C{T(0), T(0)};
#pragma unroll
    for (int f = 0; f < 3; f++) {
        T *u = f == 0 ? u0 : (f == 1 ? u1 : u2);
        const T *g = f == 0 ? g0 : (f == 1 ? g1 : g2);
        C *w = v[NV == 3 ? f : (f & 1)];
        if (NV == 2 && f == 1) load_c2r<T, N, false, EV>(v[0], h2, rA, rB, t, nullptr);  // component 2 in flight
        T ga[E], gb[E];
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int x = t + q * P;
            ga[q] = g[xA + x];
            gb[q] = g[xB + x];
        }
        fft_dist<T, N, +1, 1, EV>(w, L, t, tw);
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int x = t + q * P;
            if (u != nullptr) {
                u[xA + x] = w[q].x;
                u[xB + x] = w[q].y;
            }
            cv[q] = C{cv[q].x + w[q].x * ga[q], cv[q].y + w[q].y * gb[q]};
        }
    }
    if (part != nullptr) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < E; q++) acc += (double)cv[q].x + (double)cv[q].y;
        const double s = block_sum_f64(acc, reinterpret_cast<double *>(lds));
        if (threadIdx.x == 0) part[zb * gridDim.x + blockIdx.x] = s;
        __syncthreads();
    }
    fft_dist<T, N, -1, 1, EV>(cv, L, t, tw);
    store_r2c<T, N, EV>(cv, L, hout, rA, rB, t);
}

// P5: inverse x of (c_hat, dy c_hat, dz c_hat) -> c, grad c (real), then
// f = (lambda/eps^2)(c^3 - c) and its forward x transform (src/main.cpp:153).
template <typename T, int N>
__global__ __launch_bounds__(XT_of<T>::v) void k_x_cgrad_f(const cplx<T> *__restrict__ ch, const cplx<T> *__restrict__ dyc,
                                                  const cplx<T> *__restrict__ dzc, T *__restrict__ c,
                                                  T *__restrict__ gx, T *__restrict__ gy, T *__restrict__ gz,
                                                  cplx<T> *__restrict__ fout, T fscale, Dims d,
                                                  const double *__restrict__ kxv, const cplx<T> *__restrict__ tw,
                                                  const T *__restrict__ W)
{
    FIPI_XPASS_SETUP(T, N)
    const size_t rA = zb * d.hs + (size_t)rl * d.nxp, rB = rA + d.nxp;
    const size_t xA = zb * d.rs + (size_t)rl * d.nx, xB = xA + d.nx;
    // all inputs in flight first
    C cv[E], v[3][E];
    load_c2r<T, N, false>(cv, ch, rA, rB, t, kxv);
    load_c2r<T, N, true>(v[0], ch, rA, rB, t, kxv);
    load_c2r<T, N, false>(v[1], dyc, rA, rB, t, kxv);
    load_c2r<T, N, false>(v[2], dzc, rA, rB, t, kxv);
    fft_dist<T, N, +1>(cv, L, t, tw);
#pragma unroll
    for (int q = 0; q < E; q++) {
        const int x = t + q * P;
        c[xA + x] = cv[q].x;
        c[xB + x] = cv[q].y;
    }
#pragma unroll
    for (int f = 0; f < 3; f++) {
        T *g = f == 0 ? gx : (f == 1 ? gy : gz);
        fft_dist<T, N, +1>(v[f], L, t, tw);
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int x = t + q * P;
            g[xA + x] = v[f][q].x;
            g[xB + x] = v[f][q].y;
        }
    }
#pragma unroll
    for (int q = 0; q < E; q++) {
        const T a = cv[q].x, b = cv[q].y;
        cv[q] = C{(a * a * a - a) * fscale, (b * b * b - b) * fscale};
        if (W) {
            const int x = t + q * P;
            cv[q].x += W[xA + x];
            cv[q].y += W[xB + x];
        }
    }
    fft_dist<T, N, -1>(cv, L, t, tw);
    store_r2c<T, N>(cv, L, fout, rA, rB, t);
}

// Spread sum of one grid node: x on its line and the three force components.
template <typename T>
struct SpreadNode {
    int x;
    T v[3];
};

// Adds the spread sums of real lines rl, rl+1 to component f of a packed
// line (thread t holds elements t + q*P of both lines).  nodes[line_start[l]
// + k], k < node_cnt[l], are the per-node spread sums of real line l
// (unique x per line), each already summed in ascending particle order
// (cluster.cpp:355-407); every thread adds whole nodes, so there is one add
// per node and no ordering between threads.
template <typename T, int N>
__device__ __forceinline__ void add_spread(cplx<T> *v, cplx<T> *L, int t, int rl, int f,
                                           const int *__restrict__ line_start, const int *__restrict__ node_cnt,
                                           const SpreadNode<T> *__restrict__ nodes)
{
    constexpr int E = Line<T, N>::E, P = Line<T, N>::P;
    T *Lr = reinterpret_cast<T *>(L);  // natural-order buffer: line rl then rl+1
#pragma unroll
    for (int q = 0; q < E; q++) {
        const int x = t + q * P;
        Lr[x] = v[q].x;
        Lr[N + x] = v[q].y;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int b = line_start[rl + h], cnt = node_cnt[rl + h];
        for (int k = t; k < cnt; k += P) {
            const SpreadNode<T> nd = nodes[b + k];
            Lr[h * N + nd.x] += nd.v[f];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E; q++) {
        const int x = t + q * P;
        v[q] = cplx<T>{Lr[x], Lr[N + x]};
    }
    __syncthreads();
}

// P5 + P9: inverse x of (c_hat, dx c_hat, dy c_hat, dz c_hat, lambda k^2
// c_hat) -> c and grad c (real, stored); f = (lambda/eps^2)(c^3 - c)
// (src/main.cpp:81,153) and its forward x transform (the next step's CH
// update takes it from there); mu = f + lambda (-lap c) (src/main.cpp:157);
// the force F = mu grad c + particle spread + buoyancy
// (src/main.cpp:149-172) and its forward x transform.  The order of the
// sums is the reference's ((mu grad c + spread) + buoyancy); buoyancy
// enters the z component only: g = (0, 0, gravity) (field.cpp:122-133).
// In place: F_x over c_hat, F_y over dz c_hat, F_z over lambda k^2 c_hat.
// Each of these arrays is consumed by a transform (whose LDS barriers
// every thread of the workgroup passes) before it is overwritten.
template <typename T, int N>
__global__ __launch_bounds__(XT_of<T>::v, (sizeof(T) == 4 ? 4 : 1)) void k_x_cgrad_force(
    cplx<T> *ch /* in: c_hat (X); out: F_x (X) */, const cplx<T> *__restrict__ dyc,
    cplx<T> *dzc /* in: dz c_hat (X); out: F_y (X) */, cplx<T> *lap /* in: lambda k^2 c_hat (X); out: F_z (X) */,
    cplx<T> *__restrict__ fout, T *__restrict__ c, T *__restrict__ gx, T *__restrict__ gy, T *__restrict__ gz,
    const int *__restrict__ line_start, const int *__restrict__ node_cnt, const SpreadNode<T> *__restrict__ nodes,
    int use_spread, T fscale, const T *__restrict__ W, T rho1, T rho2, T gravity, Dims d,
    const double *__restrict__ kxv, const cplx<T> *__restrict__ tw)
{
    FIPI_XPASS_SETUP(T, N)
    FIPI_TW_TO_LDS(T, N, tw)
    const size_t rA = zb * d.hs + (size_t)rl * d.nxp, rB = rA + d.nxp;
    const size_t xA = zb * d.rs + (size_t)rl * d.nx, xB = xA + d.nx;
    // spread CSR over the lines of all boxes: box zb's lines follow box 0's
    line_start = boxp(line_start, zb, (size_t)(d.ny * d.nz));
    node_cnt = boxp(node_cnt, zb, (size_t)(d.ny * d.nz));
    // Thread-private LDS, each value written and read by the same thread (no
    // barrier): FP64 keeps the buoyancy of this thread's points, FP32 the
    // packed i kx c_hat input of the x component.  FP64 reads c_hat a second
    // time instead (the extra LDS would cost occupancy); FP32 forms the
    // buoyancy again from c, reloaded where the z component needs it (the
    // workgroup stored it, so the rows are in L2): without that array its
    // LDS and 124 VGPRs allow four waves per SIMD instead of three (the
    // pass 0.231 -> 0.212 ms, bitwise the same results).
    constexpr bool DX_LDS = sizeof(T) == 4;
    constexpr bool B_LDS = sizeof(T) == 8;
    __shared__ T bst[B_LDS ? LPW * 2 * N : 1];
    __shared__ C dst[DX_LDS ? LPW * N : 1];
    T *bq = bst + (B_LDS ? lp * 2 * N : 0);
    // mu <- lambda (-lap c), then + f(c); one array in flight per transform
    C mu[E];
    load_c2r<T, N, false>(mu, lap, rA, rB, t, nullptr);
    fft_dist<T, N, +1>(mu, L, t, tw);
    {
        C cv[E], fv[E];
        if constexpr (DX_LDS) {
            C dv[E];
            load_c2r<T, N, true>(dv, ch, rA, rB, t, kxv);
#pragma unroll
            for (int q = 0; q < E; q++) dst[lp * N + t + q * P] = dv[q];
        }
        load_c2r<T, N, false>(cv, ch, rA, rB, t, nullptr);
        fft_dist<T, N, +1>(cv, L, t, tw);
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int x = t + q * P;
            const T a = cv[q].x, b = cv[q].y;
            c[xA + x] = a;
            c[xB + x] = b;
            // ((1+c)/2*rho1 + (1-c)/2*rho2), src/main.cpp:172
            if constexpr (B_LDS) {
                bq[x] = (a + T(1)) / T(2) * rho1 + (a * T(-1) + T(1)) / T(2) * rho2;
                bq[N + x] = (b + T(1)) / T(2) * rho1 + (b * T(-1) + T(1)) / T(2) * rho2;
            }
            fv[q] = C{(a * a * a - a) * fscale, (b * b * b - b) * fscale};
            if (W) {
                fv[q].x += W[xA + x];
                fv[q].y += W[xB + x];
            }
            mu[q] = C{fv[q].x + mu[q].x, fv[q].y + mu[q].y};
        }
        fft_dist<T, N, -1>(fv, L, t, tw);
        store_r2c<T, N>(fv, L, fout, rA, rB, t);
    }
    // components in the order x, z, y: each output array is consumed first
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int f = k == 0 ? 0 : (k == 1 ? 2 : 1);
        T *g = f == 0 ? gx : (f == 1 ? gy : gz);
        C v[E];
        if (f == 0) {
            if constexpr (DX_LDS) {
#pragma unroll
                for (int q = 0; q < E; q++) v[q] = dst[lp * N + t + q * P];
            } else {
                load_c2r<T, N, true>(v, ch, rA, rB, t, kxv);
            }
        }
        else load_c2r<T, N, false>(v, f == 1 ? dyc : dzc, rA, rB, t, nullptr);
        fft_dist<T, N, +1>(v, L, t, tw);
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int x = t + q * P;
            g[xA + x] = v[q].x;
            g[xB + x] = v[q].y;
            v[q] = C{mu[q].x * v[q].x, mu[q].y * v[q].y};
        }
        if (use_spread) add_spread<T, N>(v, L, t, rl, f, line_start, node_cnt, nodes);
        if (f == 2) {
            // + buoyancy * gravity, src/main.cpp:172
#pragma unroll
            for (int q = 0; q < E; q++) {
                const int x = t + q * P;
                if constexpr (B_LDS) {
                    v[q] = C{v[q].x + bq[x] * gravity, v[q].y + bq[N + x] * gravity};
                } else {
                    // the same expression on the same stored values
                    const T a = c[xA + x], b = c[xB + x];
                    const T ba = (a + T(1)) / T(2) * rho1 + (a * T(-1) + T(1)) / T(2) * rho2;
                    const T bb = (b + T(1)) / T(2) * rho1 + (b * T(-1) + T(1)) / T(2) * rho2;
                    v[q] = C{v[q].x + ba * gravity, v[q].y + bb * gravity};
                }
            }
        }
        fft_dist<T, N, -1>(v, L, t, tw);
        store_r2c<T, N>(v, L, f == 0 ? ch : (f == 1 ? dzc : lap), rA, rB, t);
    }
}

// ---------------------------------------------------------------------------
// y and z passes: a tile of TL neighbouring kx columns; lane = column,
// t = position in the line (distribution D along y or z).
// ---------------------------------------------------------------------------
// Column tile and row (y for the z passes, z for the y passes) of a
// workgroup.  Tiles narrower than 128 bytes share each 128-byte line of a
// row with their neighbour; the GPU deals workgroups to its 8 XCDs (each
// with its own L2) round-robin in launch order, so with such tiles the grid
// has an even number of tiles per row and workgroups 16q + r and 16q + 8 + r
// -- the same XCD, one after the other there -- take the two tiles of one
// line: the line is fetched once.  Needs gridDim.x * gridDim.y to be a
// multiple of 16 (every supported edge length is a multiple of 16).
template <bool PAIR>
__device__ __forceinline__ void col_tile(int &tile, int &row)
{
    if constexpr (!PAIR) {
        tile = blockIdx.x;
        row = blockIdx.y;
    } else {
        const int np = (int)gridDim.x >> 1;
        const int l = (int)(blockIdx.x + gridDim.x * blockIdx.y);
        const int pr = 8 * (l >> 4) + (l & 7);
        tile = 2 * (pr % np) + ((l >> 3) & 1);
        row = pr / np;
    }
}

// tiles of TL columns per row for nxh columns (even when they pair)
__host__ __device__ constexpr int col_tiles(int nxh, int TL, int bytes)
{
    return bytes < 128 ? ((nxh + TL - 1) / TL + 1) / 2 * 2 : (nxh + TL - 1) / TL;
}

#define FIPI_TILE_SETUP_E(T, N, TLV, EV)                             \
    constexpr int E = Line<T, N, EV>::E, P = Line<T, N, EV>::P, LL = Line<T, N, EV>::LL; \
    constexpr int TL = TLV;                                           \
    using C = cplx<T>;                                                \
    __shared__ C lds[TL * LL];                                        \
    const int lane = threadIdx.x % TL, t = threadIdx.x / TL;          \
    int tile_, row_;                                                  \
    col_tile<(TL * (int)sizeof(C) < 128)>(tile_, row_);               \
    const int kx = tile_ * TL + lane;                                 \
    const bool ok = kx < d.nxh;                                       \
    const size_t zb = blockIdx.z;                                     \
    C *L = lds + lane;                                                \
    (void)E; (void)P;
#define FIPI_TILE_SETUP(T, N, TLV) FIPI_TILE_SETUP_E(T, N, TLV, (PlanE_for<T, N>::E))

// Values per thread of the y passes (their FFT plan).  Default: 256-point
// FP32 lines use 16 (one LDS exchange, twice the bytes in flight per
// thread; these passes hold at most three fields), everything else the
// default of its size.  Alt: the other plan of the size (0: none); a tuner
// picks between the two.
template <typename T, int N> struct YE {
    static constexpr int v = (sizeof(T) == 4 && N == 256) ? 16 : PlanE_for<T, N>::E;
    // FP64 256: 4.4.4.4 (16.16 would not fit the registers of two waves per
    // SIMD); FP32 256: 8.4.8; 64 and 128: the 4-value plans
    static constexpr int alt =
        N == 256 ? (v == 16 ? 8 : (v == 8 ? 4 : 8)) : ((N == 64 || N == 128) ? (v == 4 ? 8 : 4) : 0);
    static_assert(alt != v, "the alternative y plan must differ from the default");
};

// Element offsets within one box (the box offset is applied to the
// pointer, which stays uniform), turned into 32-bit byte offsets: a load or
// store then takes the pointer in scalar registers and one 32-bit vector
// offset, not a 64-bit vector address.  One box's array is below 4 GiB for
// every supported edge length.
template <typename C_>
__device__ __forceinline__ const C_ *at_b32(const C_ *a, uint32_t i)
{
    return reinterpret_cast<const C_ *>(reinterpret_cast<const char *>(a) + (uint32_t)(i * (uint32_t)sizeof(C_)));
}
template <typename C_>
__device__ __forceinline__ C_ *at_b32(C_ *a, uint32_t i)
{
    return reinterpret_cast<C_ *>(reinterpret_cast<char *>(a) + (uint32_t)(i * (uint32_t)sizeof(C_)));
}

// Lanes past the last column (ok false) load column 0 of their rows (every
// caller clamps base), so the loads need no branch, and zero the values.
template <typename T, int N, int EV = PlanE_for<T, N>::E>
__device__ __forceinline__ void tile_load(cplx<T> *v, const cplx<T> *__restrict__ a, uint32_t base, uint32_t st,
                                          int t, bool ok)
{
    constexpr int E = EV, P = Line<T, N, EV>::P;
#pragma unroll
    for (int q = 0; q < E; q++) {
        const cplx<T> x = *at_b32(a, base + st * (t + q * P));
        v[q] = ok ? x : cplx<T>{0, 0};
    }
}

template <typename T, int N, int EV = PlanE_for<T, N>::E>
__device__ __forceinline__ void tile_store(const cplx<T> *v, cplx<T> *__restrict__ a, uint32_t base, uint32_t st,
                                           int t, bool ok)
{
    constexpr int E = EV, P = Line<T, N, EV>::P;
    if (ok) {
#pragma unroll
        for (int q = 0; q < E; q++) *at_b32(a, base + st * (t + q * P)) = v[q];
    }
}

// y pass: transforms along y of NF fields in place.  GRAD (inverse only):
// field 0 also yields i*ky*field0 into `extra` (zero on ky = NY/2).
template <typename T, int N, int SIGN, int NF, bool GRAD, int TB, int EY>
__global__ __launch_bounds__((TileW<T, TB>::L * Line<T, N, EY>::P)) void k_y(cplx<T> *__restrict__ a0, cplx<T> *__restrict__ a1,
                                                              cplx<T> *__restrict__ a2, cplx<T> *__restrict__ extra,
                                                              Dims d, const double *__restrict__ kyv,
                                                              const cplx<T> *__restrict__ tw)
{
    FIPI_TILE_SETUP_E(T, N, (TileW<T, TB>::L), EY)
    FIPI_TW_TO_LDS(T, N, tw)
    a0 = boxp(a0, zb, d.hs);
    a1 = boxp(a1, zb, d.hs);
    a2 = boxp(a2, zb, d.hs);
    extra = boxp(extra, zb, d.hs);
    const int z = row_;
    const uint32_t base = (uint32_t)hidx(d, ok ? kx : 0, 0, z);
    const uint32_t st = d.nxp;
    auto fld = [&](int f) { return f == 0 ? a0 : (f == 1 ? a1 : a2); };
    C v[NF][E];
#pragma unroll
    for (int f = 0; f < NF; f++) tile_load<T, N, EY>(v[f], fld(f), base, st, t, ok);
#pragma unroll
    for (int f = 0; f < NF; f++) {
        C *a = fld(f);
        if (GRAD && f == 0) {
            C w[E];
#pragma unroll
            for (int q = 0; q < E; q++) {
                const int y = t + q * P;
                const T kv = y == N / 2 ? T(0) : (T)kyv[y];
                w[q] = C{-kv * v[0][q].y, kv * v[0][q].x};
            }
            fft_dist<T, N, SIGN, TL, EY>(w, L, t, tw);
            tile_store<T, N, EY>(w, extra, base, st, t, ok);
        }
        fft_dist<T, N, SIGN, TL, EY>(v[f], L, t, tw);
        tile_store<T, N, EY>(v[f], a, base, st, t, ok);
    }
}

// 1/a for a > 0 (0 for a == 0): hardware estimate refined by two
// Newton-Raphson steps r <- r + r(1 - a r), each of which squares the
// relative error; the result is within one ulp of the rounded quotient.
__device__ __forceinline__ double recip_nr(double a)
{
    double r = __builtin_amdgcn_rcp(a);
    double e = fma(-a, r, 1.0);
    r = fma(r, e, r);
    e = fma(-a, r, 1.0);
    r = fma(r, e, r);
    return a != 0.0 ? r : 0.0;
}

struct KParams {
    double dt, mobility, lambda, mu;
    double inv_g;     // 1/(nx*ny*nz), the inverse transform's normalisation
    double inv_mu_g;  // inv_g / mu
    int first;      // 1 on the first step (first-order start, main.cpp:95-114)
    int use_conv;   // switch_ns
};

// FP64 k = 0 mode (FP32-mixed path): k0[cur] = c_hat^n(0), k0[cur^1] gets
// c_hat^{n+1}(0), k0[2] = history(0); conv(0) = FP64 sum of conv (its k = 0
// transform), formed by the hub's workgroup that holds k = 0 from the x
// pass's partial sums and kept in conv0.  k0 == nullptr: the k = 0 mode is
// carried like every other.  Box b's values are at k0 + K0_STRIDE*b,
// conv0 + b and part + nb*b.
constexpr int K0_STRIDE = 4;
struct K0 {
    double *k0;
    double *conv0;       // out: the FP64 sum of conv per box
    const double *part;  // the x pass's per-workgroup FP64 sums of conv, nb per box
    int nb;
    int cur;
};

// Fixed-order FP64 sum of part[0..nb): 256 strided partial sums, then a
// tree (k_sum_strided's order for one box), by any number of threads; the
// result is valid in every thread.  s: 256 doubles of LDS.
__device__ __forceinline__ double fixed_sum256(const double *__restrict__ part, int nb, double *s)
{
    for (int v = threadIdx.x; v < 256; v += blockDim.x) {
        double a = 0;
        for (int i = v; i < nb; i += 256) a += part[i];
        s[v] = a;
    }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        for (int v = threadIdx.x; v < o; v += blockDim.x) s[v] += s[v + o];
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

// P3: CH update.  conv (XY) -> conv_hat; with c_hat^n, f_hat^n and the
// history term hist = -c^{n-1} + 2dtM k^2 f^{n-1} + 2dt conv^{n-1}:
//   c^{n+1} = (4c^n - 4dtM k^2 f^n - 4dt conv^n + hist) / (3 + 2dtM lambda k^4)
// (src/main.cpp:117-127) or on the first step (FIRST, a template argument
// so that each instance carries one formula)
//   c^1 = (c^0 - dtM k^2 f^0 - dt conv^0) / (1 + dtM lambda k^4)  (:106-113).
// Outputs c^{n+1} (K), the new history (K), and, inverse-z transformed
// (XY): c^{n+1}/G, i kz c^{n+1}/G (gradient and real-space c) and, when
// lap_out is set, lambda k^2 c^{n+1}/G, the gradient-energy part of the
// chemical potential mu = f - lambda lap c (src/main.cpp:157), which the
// force pass completes in real space.
template <typename T, int N, int TB, bool FIRST>
__global__ __launch_bounds__((TileW<T, TB>::L * Line<T, N>::P)) void k_hub_ch(
    cplx<T> *__restrict__ conv_io /* in conv XY, out c XY */, cplx<T> *__restrict__ dzc_out,
    cplx<T> *__restrict__ lap_out, const cplx<T> *__restrict__ c_now, cplx<T> *__restrict__ c_next,
    const cplx<T> *__restrict__ f_now, cplx<T> *__restrict__ hist, KParams kp, Dims d,
    const double *__restrict__ kxv, const double *__restrict__ kyv, const double *__restrict__ kzv,
    const cplx<T> *__restrict__ tw, K0 k0)
{
    FIPI_TILE_SETUP(T, N, (TileW<T, TB>::L))
    FIPI_TW_TO_LDS(T, N, tw)
    conv_io = boxp(conv_io, zb, d.hs);
    dzc_out = boxp(dzc_out, zb, d.hs);
    lap_out = boxp(lap_out, zb, d.hs);
    c_now = boxp(c_now, zb, d.hs);
    c_next = boxp(c_next, zb, d.hs);
    f_now = boxp(f_now, zb, d.hs);
    hist = boxp(hist, zb, d.hs);
    const int y = row_;
    const uint32_t base = (uint32_t)hidx(d, ok ? kx : 0, y, 0);
    const uint32_t st = (uint32_t)d.nxp * d.ny;
    // the FP64 mass mode's conv(0), in the workgroup that holds k = 0
    double cv0 = 0.0;
    if (k0.k0 != nullptr && kp.use_conv && tile_ == 0 && y == 0) {
        cv0 = fixed_sum256(k0.part + (size_t)zb * k0.nb, k0.nb, reinterpret_cast<double *>(lds));
        if (threadIdx.x == 0) k0.conv0[zb] = cv0;
    }
    C v[E], cn[E], fn[E], hh[E];
    if (kp.use_conv) tile_load<T, N>(v, conv_io, base, st, t, ok);
    tile_load<T, N>(cn, c_now, base, st, t, ok);
    tile_load<T, N>(fn, f_now, base, st, t, ok);
    tile_load<T, N>(hh, hist, base, st, t, ok);
    if (kp.use_conv) {
        fft_dist<T, N, -1, TL>(v, L, t, tw);
    } else {
#pragma unroll
        for (int q = 0; q < E; q++) v[q] = C{0, 0};
    }
    C w[E];
    const double kxx = kxv[ok ? kx : 0], kyy = kyv[y];
#pragma unroll
    for (int q = 0; q < E; q++) {
        const int z = t + q * P;
        // (always true: the per-point branch keeps one point's FP64
        // temporaries live at a time)
        if (z >= N) {
            v[q] = w[q] = C{T(0), T(0)};
            continue;
        }
        const double kzz = kzv[z];
        double k2 = 0.0;
        k2 += kxx * kxx;
        k2 += kyy * kyy;
        k2 += kzz * kzz;
        // The update is evaluated in FP64 from the stored values (FP32 in
        // the mixed path): coefficients rounded to FP32 would bias every
        // mode's growth factor by ~1e-7 per step, and the bias compounds
        // over the run; rounding only the result is unbiased.
        const double cnx = cn[q].x, cny = cn[q].y, fnx = fn[q].x, fny = fn[q].y, cvx = v[q].x, cvy = v[q].y;
        // one reciprocal of the (real) divisor per point
        C nc;
        if constexpr (FIRST) {
            const double m1 = k2 * (kp.dt * kp.mobility);
            const double rdiv = recip_nr(k2 * k2 * (kp.dt * kp.mobility * kp.lambda) + 1.0);
            double r = cnx - m1 * fnx, im = cny - m1 * fny;
            r -= cvx * kp.dt;
            im -= cvy * kp.dt;
            nc = C{(T)(r * rdiv), (T)(im * rdiv)};
        } else {
            const double m4 = k2 * (4.0 * kp.dt * kp.mobility);
            const double b4 = 4.0 * kp.dt;
            const double rdiv = recip_nr(k2 * k2 * (2.0 * kp.dt * kp.mobility * kp.lambda) + 3.0);
            double r = cnx * 4.0 - m4 * fnx, im = cny * 4.0 - m4 * fny;
            r -= cvx * b4;
            im -= cvy * b4;
            r += (double)hh[q].x;
            im += (double)hh[q].y;
            nc = C{(T)(r * rdiv), (T)(im * rdiv)};
        }
        const double m2 = k2 * (2.0 * kp.dt * kp.mobility);
        const double b2 = 2.0 * kp.dt;
        C nh = C{(T)(-cnx + m2 * fnx + cvx * b2), (T)(-cny + m2 * fny + cvy * b2)};
        const T ig = (T)kp.inv_g;
        if (k0.k0 != nullptr && ok && kx == 0 && y == 0 && z == 0) {
            // mass mode in FP64: k^2 = 0, so only c, conv and history enter
            double *kb = k0.k0 + K0_STRIDE * zb;
            const double c0 = kb[k0.cur], h0 = kb[2];
            double n0;
            if constexpr (FIRST) n0 = c0 - cv0 * kp.dt;
            else n0 = (c0 * 4.0 - cv0 * (4.0 * kp.dt) + h0) / 3.0;
            kb[k0.cur ^ 1] = n0;
            kb[2] = -c0 + cv0 * (2.0 * kp.dt);
            nc = C{(T)n0, T(0)};
            nh = C{(T)kb[2], T(0)};
            v[q] = C{(T)(n0 * kp.inv_g), T(0)};
        } else {
            v[q] = C{nc.x * ig, nc.y * ig};
        }
        if (ok) {
            const uint32_t i = base + st * z;
            *at_b32(c_next, i) = nc;
            *at_b32(hist, i) = nh;
        }
        const T kz = z == N / 2 ? T(0) : (T)kzz;
        w[q] = C{-kz * v[q].y, kz * v[q].x};
    }
    fft_dist<T, N, +1, TL>(w, L, t, tw);
    tile_store<T, N>(w, dzc_out, base, st, t, ok);
    if (lap_out != nullptr) {
        // lambda k^2 c/G from the rounded c/G: FP64 product, one rounding
#pragma unroll
        for (int q = 0; q < E; q++) {
            const double kzz = kzv[t + q * P];
            double k2 = 0.0;
            k2 += kxx * kxx;
            k2 += kyy * kyy;
            k2 += kzz * kzz;
            const double ml = k2 * kp.lambda;
            w[q] = C{(T)(ml * (double)v[q].x), (T)(ml * (double)v[q].y)};
        }
        fft_dist<T, N, +1, TL>(w, L, t, tw);
        tile_store<T, N>(w, lap_out, base, st, t, ok);
    }
    fft_dist<T, N, +1, TL>(v, L, t, tw);
    tile_store<T, N>(v, conv_io, base, st, t, ok);
}

// Initial hub: (c XY, f XY) -> c_hat^0, f_hat^0 (K) and c/G, i kz c/G (XY).
template <typename T, int N, int TB>
__global__ __launch_bounds__((TileW<T, TB>::L * Line<T, N>::P)) void k_hub_init(cplx<T> *__restrict__ c_io,
                                                                     cplx<T> *__restrict__ f_io /* out: dz c XY */,
                                                                     cplx<T> *__restrict__ c_k,
                                                                     cplx<T> *__restrict__ f_k, KParams kp, Dims d,
                                                                     const double *__restrict__ kzv,
                                                                     const cplx<T> *__restrict__ tw)
{
    FIPI_TILE_SETUP(T, N, (TileW<T, TB>::L))
    FIPI_TW_TO_LDS(T, N, tw)
    c_io = boxp(c_io, zb, d.hs);
    f_io = boxp(f_io, zb, d.hs);
    c_k = boxp(c_k, zb, d.hs);
    f_k = boxp(f_k, zb, d.hs);
    const int y = row_;
    const uint32_t base = (uint32_t)hidx(d, ok ? kx : 0, y, 0);
    const uint32_t st = (uint32_t)d.nxp * d.ny;
    C fv[E], v[E], w[E];
    tile_load<T, N>(fv, f_io, base, st, t, ok);
    tile_load<T, N>(v, c_io, base, st, t, ok);
    fft_dist<T, N, -1, TL>(fv, L, t, tw);
    tile_store<T, N>(fv, f_k, base, st, t, ok);
    fft_dist<T, N, -1, TL>(v, L, t, tw);
    tile_store<T, N>(v, c_k, base, st, t, ok);
    const T ig = (T)kp.inv_g;
#pragma unroll
    for (int q = 0; q < E; q++) {
        const int z = t + q * P;
        v[q] = C{v[q].x * ig, v[q].y * ig};
        const T kz = z == N / 2 ? T(0) : (T)kzv[z];
        w[q] = C{-kz * v[q].y, kz * v[q].x};
    }
    fft_dist<T, N, +1, TL>(v, L, t, tw);
    tile_store<T, N>(v, c_io, base, st, t, ok);
    fft_dist<T, N, +1, TL>(w, L, t, tw);
    tile_store<T, N>(w, f_io, base, st, t, ok);
}

// P11: Stokes, u_hat = (F - k (k.F)/k^2)/(mu k^2), 0 at k = 0
// (src/main.cpp:177-178, division by zero gives zero, field.cpp:1010-1019).
// The reference keeps the real part of the inverse complex transform; its
// projection is not Hermitian where a wave-number component sits on a
// Nyquist index (k_d(N/2) = +pi/dx on both sides).  The real part equals the
// inverse of the Hermitian part, which drops the k_d k_e cross terms
// between Nyquist and non-Nyquist components; that is what is formed here.
// fz (optional): f (XY) -> f_hat (K) in place, the next CH update's input,
// after the projection (its loads wait behind the projection's transforms:
// fz is not restrict-qualified).
template <typename T, int N, int TB>
__global__ __launch_bounds__((TileW<T, TB>::L * Line<T, N>::P)) void k_hub_stokes(cplx<T> *__restrict__ a0,
                                                                       cplx<T> *__restrict__ a1,
                                                                       cplx<T> *__restrict__ a2, cplx<T> *fz,
                                                                       KParams kp, Dims d,
                                                                       const double *__restrict__ kxv,
                                                                       const double *__restrict__ kyv,
                                                                       const double *__restrict__ kzv,
                                                                       const cplx<T> *__restrict__ tw)
{
    FIPI_TILE_SETUP(T, N, (TileW<T, TB>::L))
    FIPI_TW_TO_LDS(T, N, tw)
    a0 = boxp(a0, zb, d.hs);
    a1 = boxp(a1, zb, d.hs);
    a2 = boxp(a2, zb, d.hs);
    fz = boxp(fz, zb, d.hs);
    const int y = row_;
    const uint32_t base = (uint32_t)hidx(d, ok ? kx : 0, y, 0);
    const uint32_t st = (uint32_t)d.nxp * d.ny;
    C v[3][E];
    tile_load<T, N>(v[0], a0, base, st, t, ok);
    tile_load<T, N>(v[1], a1, base, st, t, ok);
    tile_load<T, N>(v[2], a2, base, st, t, ok);
    fft_dist<T, N, -1, TL>(v[0], L, t, tw);
    fft_dist<T, N, -1, TL>(v[1], L, t, tw);
    fft_dist<T, N, -1, TL>(v[2], L, t, tw);
    const int kxi = ok ? kx : 0;
    const double kxx = kxv[kxi], kyy = kyv[y];
    const bool nx_ = kxi == d.nx / 2, ny_ = y == d.ny / 2;
#pragma unroll
    for (int q = 0; q < E; q++) {
        const int z = t + q * P;
        const double kzz = kzv[z];
        const bool nz_ = z == N / 2;
        double k2 = 0.0;
        k2 += kxx * kxx;
        k2 += kyy * kyy;
        k2 += kzz * kzz;
        const double kv[3] = {kxx, kyy, kzz};
        const bool nq[3] = {nx_, ny_, nz_};
        const double F[3][2] = {{v[0][q].x, v[0][q].y}, {v[1][q].x, v[1][q].y}, {v[2][q].x, v[2][q].y}};
        // k.F split by Nyquist class (FP64 arithmetic on the stored values)
        // (selects, not branches: adding an exact zero leaves a sum unchanged)
        double sar = 0, sai = 0, sbr = 0, sbi = 0;
#pragma unroll
        for (int e = 0; e < 3; e++) {
            const double tr = kv[e] * F[e][0], ti = kv[e] * F[e][1];
            sar += nq[e] ? 0.0 : tr;
            sai += nq[e] ? 0.0 : ti;
            sbr += nq[e] ? tr : 0.0;
            sbi += nq[e] ? ti : 0.0;
        }
        // one reciprocal per point; the velocity factor is 1/(mu k^2 G).  At
        // k = 0 the projection and the velocity are zero (division by zero
        // gives zero).  The branch also keeps one point's FP64 temporaries
        // live at a time.
        if (k2 != 0.0) {
            // the reciprocal within one ulp (recip_nr): an IEEE double
            // division is a much longer instruction sequence
            const double ik2 = recip_nr(k2);
            const double cu = ik2 * kp.inv_mu_g;
#pragma unroll
            for (int e = 0; e < 3; e++) {
                const double sr = nq[e] ? sbr : sar, si = nq[e] ? sbi : sai;
                const double pr = kv[e] * sr * ik2, pi_ = kv[e] * si * ik2;
                v[e][q] = C{(T)((F[e][0] - pr) * cu), (T)((F[e][1] - pi_) * cu)};
            }
        } else {
#pragma unroll
            for (int e = 0; e < 3; e++) v[e][q] = C{T(0), T(0)};
        }
    }
    fft_dist<T, N, +1, TL>(v[0], L, t, tw);
    tile_store<T, N>(v[0], a0, base, st, t, ok);
    fft_dist<T, N, +1, TL>(v[1], L, t, tw);
    tile_store<T, N>(v[1], a1, base, st, t, ok);
    fft_dist<T, N, +1, TL>(v[2], L, t, tw);
    tile_store<T, N>(v[2], a2, base, st, t, ok);
    if (fz != nullptr) {
        __builtin_amdgcn_sched_barrier(0);
        C f[E];
        tile_load<T, N>(f, fz, base, st, t, ok);
        fft_dist<T, N, -1, TL>(f, L, t, tw);
        tile_store<T, N>(f, fz, base, st, t, ok);
    }
}

// Plain forward / inverse z transform of one field in place (unit entries).
template <typename T, int N, int SIGN, int TB>
__global__ __launch_bounds__((TileW<T, TB>::L * Line<T, N>::P)) void k_z(cplx<T> *__restrict__ a, Dims d,
                                                              const cplx<T> *__restrict__ tw)
{
    FIPI_TILE_SETUP(T, N, (TileW<T, TB>::L))
    FIPI_TW_TO_LDS(T, N, tw)
    a = boxp(a, zb, d.hs);
    const int y = row_;
    const uint32_t base = (uint32_t)hidx(d, ok ? kx : 0, y, 0);
    const uint32_t st = (uint32_t)d.nxp * d.ny;
    C v[E];
    tile_load<T, N>(v, a, base, st, t, ok);
    fft_dist<T, N, SIGN, TL>(v, L, t, tw);
    tile_store<T, N>(v, a, base, st, t, ok);
}

}  // namespace fipi
