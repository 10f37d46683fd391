// Device FFT building blocks for gfx950 (wave64, 160 KB LDS per CU).
//
// A length-N line is held by P = N/E threads with E values each, in the
// distribution D: thread t holds elements t + q*P, q < E.  The transform is
// a Stockham auto-sort sequence of radix stages whose first and last stages
// have radix E, so a line enters and leaves every transform in D.  A forward
// transform, a pointwise k-space operation and an inverse transform
// therefore chain in registers, and loads/stores in D are coalesced (t
// fastest along a contiguous line, or neighbouring lines side by side).
//
// Twiddles W_N^t come from a host-computed table (long-double accurate).
// The small DFTs use the exact constants cos/sin(2 pi j/16).
#pragma once
#include <hip/hip_runtime.h>

namespace fipi {

template <typename T> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };
template <typename T> using cplx = typename vec2<T>::type;

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by SIGN * i
template <int SIGN, typename C> __device__ __forceinline__ C mul_si(C z)
{
    if constexpr (SIGN < 0) return C{z.y, -z.x};
    else return C{-z.y, z.x};
}
// multiply by (c, SIGN*s)
template <int SIGN, typename C, typename T> __device__ __forceinline__ C mul_cs(C z, T c, T s)
{
    const T ss = SIGN < 0 ? -s : s;
    return C{z.x * c - z.y * ss, z.x * ss + z.y * c};
}

// cos/sin(2 pi j / 16), j = 0..15
struct W16 {
    static constexpr double C1 = 0.92387953251128675612818318939678828682;  // cos(pi/8)
    static constexpr double S1 = 0.38268343236508977172845998403039886676;  // sin(pi/8)
    static constexpr double R2 = 0.70710678118654752440084436210484903928;  // sqrt(1/2)
};

template <int SIGN, typename C> __device__ __forceinline__ void dft2(C &a, C &b)
{
    C t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int SIGN, typename C> __device__ __forceinline__ void dft4(C &a, C &b, C &c, C &d)
{
    C t0 = cadd(a, c), t1 = csub(a, c), t2 = cadd(b, d), t3 = mul_si<SIGN>(csub(b, d));
    a = cadd(t0, t2);
    c = csub(t0, t2);
    b = cadd(t1, t3);
    d = csub(t1, t3);
}

template <int SIGN, typename C> __device__ __forceinline__ void dft8(C *v)
{
    using T = decltype(v[0].x);
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<SIGN>(e0, e1, e2, e3);
    dft4<SIGN>(o0, o1, o2, o3);
    const T r = (T)W16::R2;
    o1 = mul_cs<SIGN>(o1, r, r);
    o2 = mul_si<SIGN>(o2);
    o3 = mul_cs<SIGN>(o3, -r, r);
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

template <int SIGN, typename C> __device__ __forceinline__ void dft16(C *v)
{
    using T = decltype(v[0].x);
    // n = 4 n1 + n2; stage 1 over n1 for each n2
    C a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
        C p = v[n2], q = v[4 + n2], r = v[8 + n2], s = v[12 + n2];
        dft4<SIGN>(p, q, r, s);
        a[n2][0] = p; a[n2][1] = q; a[n2][2] = r; a[n2][3] = s;
    }
    const T c1 = (T)W16::C1, s1 = (T)W16::S1, r2 = (T)W16::R2;
    // twiddles W16^{n2 k1}
    a[1][1] = mul_cs<SIGN>(a[1][1], c1, s1);
    a[1][2] = mul_cs<SIGN>(a[1][2], r2, r2);
    a[1][3] = mul_cs<SIGN>(a[1][3], s1, c1);
    a[2][1] = mul_cs<SIGN>(a[2][1], r2, r2);
    a[2][2] = mul_si<SIGN>(a[2][2]);
    a[2][3] = mul_cs<SIGN>(a[2][3], -r2, r2);
    a[3][1] = mul_cs<SIGN>(a[3][1], s1, c1);
    a[3][2] = mul_cs<SIGN>(a[3][2], -r2, r2);
    a[3][3] = mul_cs<SIGN>(a[3][3], -c1, -s1);
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        C p = a[0][k1], q = a[1][k1], r = a[2][k1], s = a[3][k1];
        dft4<SIGN>(p, q, r, s);
        v[k1] = p; v[k1 + 4] = q; v[k1 + 8] = r; v[k1 + 12] = s;
    }
}

// cos/sin(2 pi k / R) for the odd-radix stages (closed forms: sqrt(3)/2,
// (sqrt(5) - 1)/4, sqrt(10 + 2 sqrt(5))/4, -(sqrt(5) + 1)/4,
// sqrt(10 - 2 sqrt(5))/4), k = 0..R-1
template <int R> struct WR;
template <> struct WR<3> {
    static constexpr double c[3] = {1.0, -0.5, -0.5};
    static constexpr double s[3] = {0.0, 0.86602540378443864676372317075293618347,
                                    -0.86602540378443864676372317075293618347};
};
template <> struct WR<5> {
    static constexpr double c[5] = {1.0, 0.30901699437494742410229341718281905886,
                                    -0.80901699437494742410229341718281905886,
                                    -0.80901699437494742410229341718281905886,
                                    0.30901699437494742410229341718281905886};
    static constexpr double s[5] = {0.0, 0.95105651629515357211643933337938214341,
                                    0.58778525229247312916870595463907276860,
                                    -0.58778525229247312916870595463907276860,
                                    -0.95105651629515357211643933337938214341};
};
template <> struct WR<6> {
    static constexpr double c[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
    static constexpr double s[6] = {0.0, 0.86602540378443864676372317075293618347,
                                    0.86602540378443864676372317075293618347, 0.0,
                                    -0.86602540378443864676372317075293618347,
                                    -0.86602540378443864676372317075293618347};
};

// direct DFT of radix R: out[m] = sum_q v[q] W_R^(SIGN q m)
template <int R, int SIGN, typename C> __device__ __forceinline__ void dft_direct(C *v)
{
    using T = decltype(v[0].x);
    C o[R];
#pragma unroll
    for (int m = 0; m < R; m++) {
        C a = v[0];
#pragma unroll
        for (int q = 1; q < R; q++) {
            const int k = (q * m) % R;
            a = cadd(a, k == 0 ? v[q] : mul_cs<SIGN>(v[q], (T)WR<R>::c[k], (T)WR<R>::s[k]));
        }
        o[m] = a;
    }
#pragma unroll
    for (int m = 0; m < R; m++) v[m] = o[m];
}

template <int R, int SIGN, typename C> __device__ __forceinline__ void dft(C *v)
{
    if constexpr (R == 3 || R == 5 || R == 6) {
        dft_direct<R, SIGN>(v);
    } else if constexpr (R == 1) {
    } else if constexpr (R == 2) {
        dft2<SIGN>(v[0], v[1]);
    } else if constexpr (R == 4) {
        dft4<SIGN>(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8) {
        dft8<SIGN>(v);
    } else if constexpr (R == 16) {
        dft16<SIGN>(v);
    } else {
        static_assert(R == 16, "unsupported radix");
    }
}

// ---------------------------------------------------------------------------
// Line transforms with E values per thread.
//
// Stockham stage s with radix R and span Ns (Govindaraju et al., "High
// performance discrete Fourier transforms on graphics processors", SC'08)
// takes butterfly j from x[j + q N/R], multiplies by W_{Ns R}^{(j mod Ns) q},
// runs a DFT-R and writes y[(j/Ns) Ns R + (j mod Ns) + q Ns].  With first
// and last radix E the first stage reads D and the last leaves the result in
// D.  Middle stages exchange through LDS.
//
// LDS slot of element i of a line: lpad(i) * S, with lpad(i) = i + i/E (one
// spare slot per E values, so the stride-E writes of the first stage spread
// over the banks; i + i/16 for E = 4 and S = 1, see lpad) and S lines
// interleaved slot by slot (the y and z tiles
// put their S = 128-byte-wide columns side by side, so a wave's accesses
// are runs of S consecutive values).
//
// Plans (E.M.E): 256 = 8.4.8, 16.16 or 4.4.4.4 (two middle stages),
// 128 = 8.2.8 or 4.8.4, 96 = 4.6.4,
// 80 = 4.5.4, 64 = 8.8 or 4.4.4, 48 = 4.3.4, 32 = 4.2.4, 16 = 4.4.  When M does not
// divide E the N/M butterflies of the middle stage are spread over the P
// threads unevenly (some threads take one more).  Lines of 256 use E = 8
// by default (fewer registers for kernels that keep three fields in
// flight); a kernel may choose another E (the FP32 y passes use 16: one LDS
// exchange instead of two, twice the bytes in flight per thread).
// ---------------------------------------------------------------------------
template <int N, int E> struct PlanE;
template <> struct PlanE<16, 4> { static constexpr int M = 0; };
template <> struct PlanE<32, 4> { static constexpr int M = 2; };
template <> struct PlanE<48, 4> { static constexpr int M = 3; };
template <> struct PlanE<64, 8> { static constexpr int M = 0; };
template <> struct PlanE<64, 4> { static constexpr int M = 4; };
template <> struct PlanE<80, 4> { static constexpr int M = 5; };
template <> struct PlanE<96, 4> { static constexpr int M = 6; };
template <> struct PlanE<128, 8> { static constexpr int M = 2; };
template <> struct PlanE<128, 4> { static constexpr int M = 8; };
template <> struct PlanE<256, 8> { static constexpr int M = 4; };
template <> struct PlanE<256, 16> { static constexpr int M = 0; };
template <> struct PlanE<256, 4> { static constexpr int M = 4; };
// radix of a second middle stage (0: none): 256 = 4.4.4.4
template <int N, int E> struct PlanM2 { static constexpr int v = 0; };
template <> struct PlanM2<256, 4> { static constexpr int v = 4; };

template <typename T, int N> struct PlanE_for {
    static constexpr int E = (N == 64 || N == 128 || N == 256) ? 8 : 4;
};

// EV: values per thread of a kernel that does not use the default
template <typename T, int N, int EV = PlanE_for<T, N>::E> struct Line {
    static constexpr int E = EV;                   // values per thread
    static constexpr int M = PlanE<N, E>::M;       // middle radix (0: none)
    static constexpr int M2 = PlanM2<N, E>::v;     // second middle radix (0: none)
    static constexpr int P = N / E;                // threads per line
    static constexpr int LL = N + N / E;           // padded LDS slots per line
};

// Lines held 4 values per thread and one per LDS slot run (S = 1, the x
// passes) pad one slot per 16 instead of per 4: fewer bank conflicts in the
// 4.4.4.4 plan's exchanges (measured: the FP64 convection pass 0.339 ->
// 0.317 ms).  Slots stay below LL = N + N/E either way.
template <int E, int S = 1> __device__ __forceinline__ int lpad(int i)
{
    return i + i / ((S == 1 && E == 4) ? 16 : E);
}

// W_N^(SIGN idx) from the forward table (conjugated for SIGN > 0)
template <int SIGN, typename C>
__device__ __forceinline__ C tw_at(const C *__restrict__ tw, int idx)
{
    C w = tw[idx];
    if constexpr (SIGN > 0) w.y = -w.y;
    return w;
}

// middle Stockham stage of radix RM and span NS (the product of the
// radices before it)
template <typename T, int N, int SIGN, int S, int EV, int RM = Line<T, N, EV>::M, int NS = EV>
__device__ __forceinline__ void mid_stage(cplx<T> *lds, int t, const cplx<T> *__restrict__ tw)
{
    constexpr int E = EV, P = Line<T, N, EV>::P;
    constexpr int B = N / RM;               // butterflies of the stage
    constexpr int NB = (B + P - 1) / P;     // per thread (the last round may be partial)
    using C = cplx<T>;
    C u[NB * RM];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = t + b * P;
        if (B % P == 0 || j < B) {
#pragma unroll
            for (int q = 0; q < RM; q++) u[b * RM + q] = lds[lpad<E, S>(j + q * (N / RM)) * S];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = t + b * P;
        if (B % P == 0 || j < B) {
            const int k = j % NS;
#pragma unroll
            for (int q = 1; q < RM; q++)
                u[b * RM + q] = cmul(u[b * RM + q], tw_at<SIGN>(tw, k * q * (N / (NS * RM))));
            dft<RM, SIGN>(u + b * RM);
            const int base = (j / NS) * NS * RM + k;
#pragma unroll
            for (int q = 0; q < RM; q++) lds[lpad<E, S>(base + q * NS) * S] = u[b * RM + q];
        }
    }
    __syncthreads();
}

// In-place transform of one line held in D by threads t < P (every thread of
// the workgroup must call it: it contains barriers).  lds: slot 0 of this
// line; slots are S values apart.
template <typename T, int N, int SIGN, int S = 1, int EV = PlanE_for<T, N>::E>
__device__ __forceinline__ void fft_dist(cplx<T> *v, cplx<T> *lds, int t, const cplx<T> *__restrict__ tw)
{
    constexpr int E = EV, P = Line<T, N, EV>::P;
    dft<E, SIGN>(v);
#pragma unroll
    for (int q = 0; q < E; q++) lds[lpad<E, S>(t * E + q) * S] = v[q];
    __syncthreads();
    if constexpr (Line<T, N, EV>::M > 0) mid_stage<T, N, SIGN, S, EV>(lds, t, tw);
    if constexpr (Line<T, N, EV>::M2 > 0)
        mid_stage<T, N, SIGN, S, EV, Line<T, N, EV>::M2, EV * Line<T, N, EV>::M>(lds, t, tw);
#pragma unroll
    for (int q = 0; q < E; q++) v[q] = lds[lpad<E, S>(t + q * P) * S];
#pragma unroll
    for (int q = 1; q < E; q++) v[q] = cmul(v[q], tw_at<SIGN>(tw, t * q));
    dft<E, SIGN>(v);
    __syncthreads();
}

}  // namespace fipi
