// Particle kernels: cell list, 27-cell contact forces, trilinear
// interpolation, capillary attachment force, Heun advection and the
// deterministic spreading of particle forces to the grid.
//
// All particle quantities are FP64 in both precision modes.  Every kernel
// is deterministic: integer atomics only fix counts and slots, and every
// floating-point sum runs in a fixed order (no float atomics).
//
// Floating-point expressions follow the reference's association order and
// this file is compiled with -ffp-contract=off, so for identical inputs the
// cell membership, contact-pair set and contact forces are bitwise those of
// the reference restatement.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "field_kernels.hpp"

namespace fipi {

struct PartParams {
    int n;
    int nx, ny, nz;              // grid
    double dx, dy, dz;           // grid spacing L/N (head.hpp:37-39)
    double lx, ly, lz;
    int ncx, ncy, ncz;           // contact cells (cluster.cpp:26-28)
    double dcx, dcy, dcz;        // contact cell size (cluster.cpp:30-32)
    double margin;
    double r_p, r_eq, k_n;
    double pi, mu, dt, epsilon, sigma, a_pi, tanh2;
    double sqrt2;                // pow(2.0, 0.5) as the reference's sqrt override
    int has_ctheta;
    // the reference's unused physics: uf from u (0 off, 1 interpolate,
    // 2 interpolate2) and the Brownian force amplitude sqrt(12 pi r_p mu
    // kt/dt) (0 off) with its draw seed
    int interp_u;
    double brown_amp;
    uint64_t noise_seed;
    // slab of nbox equal boxes: contact cell and grid line indices are
    // global, box*ncell_box + cell and box*nlines_box + line; box b's
    // grid arrays start b*rs elements in
    int nbox, ncell_box, nlines_box;
    size_t rs;
};

struct PartState {
    double *p[3], *po[3], *u[3], *uo[3], *fpp[3], *fpi[3];
    double *rad, *cth;
    double *cloc, *dcdx[3];
    double *uf[3], *frand[3];    // local fluid velocity and Brownian force of the last stage
    double *tau_i;               // per-particle packing-energy contributions
    // per-particle blocks of the step's particle stage (particle_stage.hpp):
    // Carry<T> (stage 1 to stage 2) and StepOut (the step's results, from
    // which u, u_old, p_old, f_pp, f_pi, grad c, c and tau above are formed
    // on demand)
    void *carry, *out;
    double *sp[3];               // positions in cell-sorted order (sp[d][s] = p[d][ids[s]])
    int *cell;                   // contact cell of each particle
    int *err;                    // device error flags (bit 0: sorting range)
    int *mark;                   // it_mark (argmax y), persists across sorts
    int *mark_on;                // the stage-2 tau gate: f_pi of it_mark non-zero
    const int *box;              // box of each particle (null: all in box 0)
};

__device__ __forceinline__ int part_box(const int *box, int a) { return box ? box[a] : 0; }

// ---------------------------------------------------------------------------
// wave64 inclusive scan of one int per lane
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// ---------------------------------------------------------------------------
// Multi-block exclusive scan: k_scan_blocks (each block of 256 threads scans
// SCAN_ITEMS consecutive counts, 4 per thread, and writes its total),
// k_scan_finish (adds each block's offset, writes start[n]).
// ---------------------------------------------------------------------------
constexpr int SCAN_ITEMS = 1024;

__device__ __forceinline__ int block_excl_scan256(int v, int *wsum, int *total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int inc = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wid; w++) off += wsum[w];
    *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return off + inc - v;
}

__global__ __launch_bounds__(256) void k_scan_blocks(const int *__restrict__ cnt, int *__restrict__ start, int n,
                                                     int *__restrict__ bsum)
{
    __shared__ int wsum[4];
    const int b0 = blockIdx.x * SCAN_ITEMS + threadIdx.x * 4;
    int v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = (b0 + q < n) ? cnt[b0 + q] : 0;
    const int s = v[0] + v[1] + v[2] + v[3];
    int tot;
    int run = block_excl_scan256(s, wsum, &tot);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (b0 + q < n) start[b0 + q] = run;
        run += v[q];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Each block adds the sum of the block totals before it (bsum[0..b), a
// block reduction; integer sums, so the order is immaterial); the last
// block writes start[n] = grand total.  Block 0 also reduces the argmax-y
// partials (y, index) of nparts blocks into *mark: the strictly largest
// y > 0, first index on ties (cluster.cpp:128-138); the previous value is
// kept if no y > 0.  The (y descending, index ascending) order is total,
// so the tree is exact.
__global__ __launch_bounds__(256) void k_scan_finish(int *__restrict__ start, int n, const int *__restrict__ bsum,
                                                     int nb, const double *__restrict__ ymax,
                                                     const int *__restrict__ yidx, int nparts, int *__restrict__ mark)
{
    __shared__ int ws[256];
    __shared__ double sy[256];
    __shared__ int si[256];
    const int b = blockIdx.x;
    int acc = 0;
    for (int i = threadIdx.x; i < b; i += 256) acc += bsum[i];
    ws[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) ws[threadIdx.x] += ws[threadIdx.x + o];
        __syncthreads();
    }
    const int off = ws[0];
    const int i0 = b * SCAN_ITEMS + threadIdx.x * 4;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (i0 + q < n) start[i0 + q] += off;
    if (b == nb - 1 && threadIdx.x == 0) start[n] = off + bsum[b];
    if (mark == nullptr || b != 0) return;
    double by = 0.0;
    int bi = -1;
    for (int k = threadIdx.x; k < nparts; k += 256) {
        const double oy = ymax[k];
        const int oi = yidx[k];
        if (oi >= 0 && (bi < 0 || oy > by || (oy == by && oi < bi))) { by = oy; bi = oi; }
    }
    sy[threadIdx.x] = by;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const double oy = sy[threadIdx.x + o];
            const int oi = si[threadIdx.x + o];
            if (oi >= 0 && (si[threadIdx.x] < 0 || oy > sy[threadIdx.x] ||
                            (oy == sy[threadIdx.x] && oi < si[threadIdx.x]))) {
                sy[threadIdx.x] = oy;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && si[0] >= 0) *mark = si[0];
}

// Block partial of argmax y (strictly greater, first index; y_max starts at
// 0.0, cluster.cpp:128-138) over the particles of this block (256 threads):
// a butterfly over each wave, then the four wave results in LDS.  The
// comparison is a total order on (y, -index) with -1 for "none", so the
// result does not depend on the order of the combines.
__device__ __forceinline__ void argmax_take(double &by, int &bi, double oy, int oi)
{
    if (oi >= 0 && (bi < 0 || oy > by || (oy == by && oi < bi))) {
        by = oy;
        bi = oi;
    }
}

__device__ __forceinline__ void block_argmax_y(double y, int idx, double *ymax, int *yidx, int slot = -1)
{
    __shared__ double sy[4];
    __shared__ int si[4];
    double by = y > 0.0 ? y : 0.0;
    int bi = y > 0.0 ? idx : -1;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) argmax_take(by, bi, __shfl_xor(by, o), __shfl_xor(bi, o));
    if ((threadIdx.x & 63) == 0) {
        sy[threadIdx.x >> 6] = by;
        si[threadIdx.x >> 6] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) argmax_take(by, bi, sy[w], si[w]);
        const int k = slot >= 0 ? slot : (int)blockIdx.x;
        ymax[k] = by;
        yidx[k] = bi;
    }
}

// cluster.cpp:140-142 / 197-199
__device__ __forceinline__ int cell_coord(double x, double dc, double margin)
{
    return (int)floor(x / dc + margin / dc);
}

__device__ __forceinline__ int cell_of(const PartParams &P, double x, double y, double z, int *ci, int *cj,
                                       int *ck)
{
    const int i = cell_coord(x, P.dcx, P.margin), j = cell_coord(y, P.dcy, P.margin),
              k = cell_coord(z, P.dcz, P.margin);
    *ci = i; *cj = j; *ck = k;
    if (i < 0 || i >= P.ncx || j < 0 || j >= P.ncy || k < 0 || k >= P.ncz) return -1;
    return i + j * P.ncx + k * P.ncx * P.ncy;
}

// Histogram of contact cells; flags particles outside the sorting range.
__global__ __launch_bounds__(256) void k_cell_count(PartParams P, const double *__restrict__ px,
                                                    const double *__restrict__ py, const double *__restrict__ pz,
                                                    const int *__restrict__ box, int *__restrict__ cell,
                                                    int *__restrict__ cnt, int *__restrict__ err,
                                                    double *__restrict__ ymax, int *__restrict__ yidx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    block_argmax_y(i < P.n ? py[i] : -1.0, i, ymax, yidx);
    if (i >= P.n) return;
    int a, b, c;
    const int cc = cell_of(P, px[i], py[i], pz[i], &a, &b, &c);
    if (cc < 0) { cell[i] = -1; atomicOr(err, 1); return; }
    const int g = cc + part_box(box, i) * P.ncell_box;
    cell[i] = g;
    atomicAdd(&cnt[g], 1);
}

__global__ void k_cell_scatter(int n, const int *__restrict__ cell, const int *__restrict__ start,
                               int *__restrict__ cursor, int *__restrict__ ids)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int cc = cell[i];
    if (cc < 0) return;
    const int slot = atomicAdd(&cursor[cc], 1);
    ids[start[cc] + slot] = i;
}

// Ascending ids within each cell (list::append order, linked_list.cpp:30-43);
// the positions are gathered in the same order (sp, may be null).
// cnt and cursor of the cell are cleared for the next build.
__global__ void k_cell_sort(int ncell, const int *__restrict__ start, int *__restrict__ ids, PartState S,
                            int *__restrict__ cnt, int *__restrict__ cursor)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    cnt[c] = 0;
    cursor[c] = 0;
    const int b = start[c], e = start[c + 1];
    for (int i = b + 1; i < e; i++) {
        const int v = ids[i];
        int j = i - 1;
        while (j >= b && ids[j] > v) { ids[j + 1] = ids[j]; j--; }
        ids[j + 1] = v;
    }
    if (S.sp[0]) {
        for (int i = b; i < e; i++) {
            const int a = ids[i];
            S.sp[0][i] = S.p[0][a];
            S.sp[1][i] = S.p[1][a];
            S.sp[2][i] = S.p[2][a];
        }
    }
}

// ---------------------------------------------------------------------------
// Contact forces (cluster.cpp:174-275)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Deterministic stand-in for the rand() jitter of coincident particles
// (cluster.cpp:234-242), for the ordered pair i < j.
__device__ __forceinline__ double jitter(uint32_t i, uint32_t j, int a, double r_p)
{
    uint64_t h = splitmix64(((uint64_t)i << 32) | (uint64_t)j);
    h = splitmix64(h + (uint64_t)(a + 1));
    return (double)(h % 1000u) / 1e3 * 2.0 * r_p - r_p;
}

__device__ __forceinline__ double vnorm(double x, double y, double z)
{
    // primitive.hpp:58, pow(pow(x,2)+pow(y,2)+pow(z,2),0.5)
    return sqrt((x * x + y * y) + z * z);
}

// d = p_a - p_b with the reference's minimum image (cluster.cpp:223-232);
// coincident particles get the jitter of the ordered pair (lo, hi), negated
// when a > b so that F(a,b) = -F(b,a) bitwise.
__device__ __forceinline__ void pair_disp(const PartParams &P, const PartState &S, int a, int b, double *d)
{
    d[0] = S.p[0][a] - S.p[0][b];
    d[1] = S.p[1][a] - S.p[1][b];
    d[2] = S.p[2][a] - S.p[2][b];
    const double L[3] = {P.lx, P.ly, P.lz};
    const double th[3] = {P.dcx * (P.ncx - 2), P.dcy * (P.ncy - 2), P.dcz * (P.ncz - 2)};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        if (d[q] > th[q]) d[q] -= L[q];
        else if (d[q] < -th[q]) d[q] += L[q];
    }
    if (vnorm(d[0], d[1], d[2]) == 0.0) {
        const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double jv = jitter(lo, hi, q, P.r_p);
            d[q] = a < b ? jv : -jv;
        }
    }
}

// Force on a from b if in contact; returns false otherwise.
__device__ __forceinline__ bool pair_force(const PartParams &P, const PartState &S, int a, int b, double rmp,
                                           double *f, double *d)
{
    pair_disp(P, S, a, b, d);
    const double nrm = vnorm(d[0], d[1], d[2]);
    if (!(nrm < P.r_eq)) return false;
    const double s = (P.r_eq - nrm) * rmp * P.k_n;
    f[0] = (d[0] / nrm) * s;
    f[1] = (d[1] / nrm) * s;
    f[2] = (d[2] / nrm) * s;
    return true;
}

__device__ __forceinline__ int wrapc(int v, int n) { return (v + n) % n; }

// Reference-order force sum by repeated scans, for the waves whose partner
// lists do not fit the ordered path of k_contact_par (kept exact).
__device__ void contact_slow(const PartParams &P, const PartState &S, const int *__restrict__ start,
                             const int *__restrict__ ids, int a, int o, int pp, int q, int cb, double rmp,
                             bool tau_on, double *fo, double *tau_out)
{
    double fx = 0.0, fy = 0.0, fz = 0.0;
    int last = -1;
    for (;;) {
        int best = 0x7fffffff;
        for (int k = q - 1; k <= q + 1; k++)
            for (int j = pp - 1; j <= pp + 1; j++)
                for (int i = o - 1; i <= o + 1; i++) {
                    const int co = cb + wrapc(i, P.ncx) + wrapc(j, P.ncy) * P.ncx + wrapc(k, P.ncz) * P.ncx * P.ncy;
                    for (int e = start[co]; e < start[co + 1]; e++) {
                        const int b = ids[e];
                        if (b >= a || b <= last || b >= best) continue;
                        double d[3];
                        pair_disp(P, S, a, b, d);
                        if (vnorm(d[0], d[1], d[2]) < P.r_eq) best = b;
                    }
                }
        if (best == 0x7fffffff) break;
        double f[3], d[3];
        pair_force(P, S, a, best, rmp, f, d);
        fx = fx + f[0];
        fy = fy + f[1];
        fz = fz + f[2];
        last = best;
    }
    double tau = 0.0;
    for (int k = q - 1; k <= q + 1; k++)
        for (int j = pp - 1; j <= pp + 1; j++)
            for (int i = o - 1; i <= o + 1; i++) {
                const int co = cb + wrapc(i, P.ncx) + wrapc(j, P.ncy) * P.ncx + wrapc(k, P.ncz) * P.ncx * P.ncy;
                for (int e = start[co]; e < start[co + 1]; e++) {
                    const int b = ids[e];
                    if (b <= a) continue;
                    double f[3], d[3];
                    if (pair_force(P, S, a, b, rmp, f, d)) {
                        fx = fx + f[0];
                        fy = fy + f[1];
                        fz = fz + f[2];
                        if (tau_on) tau += 0.5 * ((f[0] * d[0] + f[1] * d[1]) + f[2] * d[2]);
                    }
                }
            }
    fo[0] = fx;
    fo[1] = fy;
    fo[2] = fz;
    *tau_out = tau;
}

// rmp: the contact ramp; when t_dev is set it is ramp(*t_dev + 1, ramp_t)
// (primitive.cpp:122-129) from the device step counter instead.
__device__ __forceinline__ double ramp_dev(double t, double thr)
{
    if (t < 0) return 0.0;
    if (t < thr) return t / thr;
    return 1.0;
}

// Contact forces (cluster.cpp:174-275) with LANES lanes per particle
// (256/LANES particles per workgroup, taken in cell-sorted order): lane c
// scans neighbour cells c*CELLS .. c*CELLS+CELLS-1 in the reference's
// traversal order (k outer, then j, then i) over the cell-sorted positions,
// so the lanes' partners, concatenated in lane order, are in traversal
// order; an ordered prefix sum over the lanes places the two partner lists
// in LDS.  Each lane writes the pair forces of its own partners next to
// their ids, so the force evaluations run in parallel.  The first lane sums
// f_pp[a] from LDS in the order the reference accumulates it: the
// contributions of partners b < a in ascending b (added when the outer loop
// visited b), then partners b > a in traversal order, with F(a,b) = -F(b,a)
// bitwise.  tau_i[a] = sum over b > a of 0.5 F.d when f_pi[a] and
// f_pi[mark] are non-zero (cluster.cpp:256-259).  A wave in which some lane
// holds more than CP_HITS partners on one side, or some particle more than
// NB, takes the exact repeated-scan path instead.  Few particles (small
// boxes) take 32 lanes, one per neighbour cell, for the shortest chain of
// dependent loads; many take 8 lanes of 4 cells each, a quarter of the
// waves, which leaves more of the GPU to the field passes beside them.
constexpr int CP_HITS = 6;                   // partners per lane and side
constexpr int CP_WIDE_MIN = 16384;           // particles from which 8 lanes are used
template <int LANES> struct ContactShape {
    static constexpr int CELLS = (27 + LANES - 1) / LANES;  // neighbour cells per lane
    static constexpr int NB = LANES > 16 ? LANES : 16;      // partners per side in LDS
    static constexpr int PPB = 256 / LANES;                 // particles per workgroup
    static_assert(NB % LANES == 0, "rank pass: whole entries per lane");
};

__device__ __forceinline__ void pair_disp_pos(const PartParams &P, double ax, double ay, double az, double bx,
                                              double by, double bz, int a, int b, double *d)
{
    d[0] = ax - bx;
    d[1] = ay - by;
    d[2] = az - bz;
    const double L[3] = {P.lx, P.ly, P.lz};
    const double th[3] = {P.dcx * (P.ncx - 2), P.dcy * (P.ncy - 2), P.dcz * (P.ncz - 2)};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        if (d[q] > th[q]) d[q] -= L[q];
        else if (d[q] < -th[q]) d[q] += L[q];
    }
    if (vnorm(d[0], d[1], d[2]) == 0.0) {
        const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double jv = jitter(lo, hi, q, P.r_p);
            d[q] = a < b ? jv : -jv;
        }
    }
}

// inclusive sum over the LANES lanes of each particle
template <int LANES>
__device__ __forceinline__ int seg_incl_scan(int v)
{
    const int lane = threadIdx.x & (LANES - 1);
#pragma unroll
    for (int o = 1; o < LANES; o <<= 1) {
        const int u = __shfl_up(v, o, LANES);
        if (lane >= o) v += u;
    }
    return v;
}

// Pair force on a from the partner at cell-sorted slot e, as pair_force()
// computes it from the same positions; w[3] = 0.5 F.d (the torque term).
__device__ __forceinline__ void contact_entry(const PartParams &P, const PartState &S, double ax, double ay,
                                              double az, int a, int b, int e, double rmp, double *w)
{
    double d[3];
    pair_disp_pos(P, ax, ay, az, S.sp[0][e], S.sp[1][e], S.sp[2][e], a, b, d);
    const double nrm = vnorm(d[0], d[1], d[2]);
    const double sc = (P.r_eq - nrm) * rmp * P.k_n;
    const double f0 = (d[0] / nrm) * sc, f1 = (d[1] / nrm) * sc, f2 = (d[2] / nrm) * sc;
    w[0] = f0;
    w[1] = f1;
    w[2] = f2;
    w[3] = 0.5 * ((f0 * d[0] + f1 * d[1]) + f2 * d[2]);
}

template <int LANES>
__global__ __launch_bounds__(256) void k_contact_par(PartParams P, PartState S, const int *__restrict__ start,
                                                     const int *__restrict__ ids, double rmp,
                                                     const int *__restrict__ t_dev, double ramp_t,
                                                     int *__restrict__ overflow)
{
    using CS = ContactShape<LANES>;
    constexpr int CELLS = CS::CELLS, NB = CS::NB, PPB = CS::PPB;
    __shared__ int lists[PPB][2][NB];
    __shared__ double forces[PPB][2][NB][4];  // per partner: F(a,b) and 0.5 F.d
    const int sub = threadIdx.x / LANES, c = threadIdx.x % LANES;
    const int s = blockIdx.x * PPB + sub;
    const int nsorted = start[P.ncell_box * P.nbox];
    const bool live = s < P.n && s < nsorted;
    if (t_dev) rmp = ramp_dev((double)(*t_dev + 1), ramp_t);
    int a = 0, nl = 0, nu = 0;
    int hl[CP_HITS], hu[CP_HITS];  // partner ids
    int sl[CP_HITS], su[CP_HITS];  // their cell-sorted slots
    bool ovf = false;
    double ax = 0, ay = 0, az = 0;
    int o = 0, pp = 0, q = 0, cb = 0;
    if (live) {
        a = ids[s];
        ax = S.sp[0][s];
        ay = S.sp[1][s];
        az = S.sp[2][s];
        cb = part_box(S.box, a) * P.ncell_box;
        cell_of(P, ax, ay, az, &o, &pp, &q);
#pragma unroll
        for (int r = 0; r < CELLS; r++) {
            const int cc = c * CELLS + r;
            if (cc >= 27) break;
            const int k = q - 1 + cc / 9, j = pp - 1 + (cc / 3) % 3, i = o - 1 + cc % 3;
            const int co = cb + wrapc(i, P.ncx) + wrapc(j, P.ncy) * P.ncx + wrapc(k, P.ncz) * P.ncx * P.ncy;
            const int e1 = start[co + 1];
            for (int e = start[co]; e < e1; e++) {
                const int b = ids[e];
                if (b == a) continue;
                double d[3];
                pair_disp_pos(P, ax, ay, az, S.sp[0][e], S.sp[1][e], S.sp[2][e], a, b, d);
                if (vnorm(d[0], d[1], d[2]) < P.r_eq) {
                    if (b < a) {
                        if (nl < CP_HITS) { hl[nl] = b; sl[nl] = e; }
                        else ovf = true;
                        nl++;
                    } else {
                        if (nu < CP_HITS) { hu[nu] = b; su[nu] = e; }
                        else ovf = true;
                        nu++;
                    }
                }
            }
        }
    }
    // ordered compaction of the lanes' partners and their forces into LDS
    const int il = seg_incl_scan<LANES>(nl), iu = seg_incl_scan<LANES>(nu);
    const int tl = __shfl(il, LANES - 1, LANES), tu = __shfl(iu, LANES - 1, LANES);
    const bool any_ovf = __any(ovf) != 0;  // per wave: all its particles fall back
    const bool fits = !any_ovf && tl <= NB && tu <= NB;
    if (fits) {
#pragma unroll
        for (int x = 0; x < CP_HITS; x++)
            if (x < nl) {
                lists[sub][0][il - nl + x] = hl[x];
                contact_entry(P, S, ax, ay, az, a, hl[x], sl[x], rmp, forces[sub][0][il - nl + x]);
            }
#pragma unroll
        for (int x = 0; x < CP_HITS; x++)
            if (x < nu) {
                lists[sub][1][iu - nu + x] = hu[x];
                contact_entry(P, S, ax, ay, az, a, hu[x], su[x], rmp, forces[sub][1][iu - nu + x]);
            }
    }
    __syncthreads();
    // partners b < a into ascending id: entry x moves to its stable rank (a
    // partner seen through two wrapped neighbour cells appears twice, with
    // equal forces)
    constexpr int PER = NB / LANES;
    int rk[PER];
    double w[PER][3];
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const int x = c + t * LANES;
        rk[t] = -1;
        w[t][0] = w[t][1] = w[t][2] = 0.0;
        if (fits && x < tl) {
            const int v = lists[sub][0][x];
            int r = 0;
            for (int y = 0; y < tl; y++) {
                const int u = lists[sub][0][y];
                r += (u < v || (u == v && y < x)) ? 1 : 0;
            }
            rk[t] = r;
            w[t][0] = forces[sub][0][x][0];
            w[t][1] = forces[sub][0][x][1];
            w[t][2] = forces[sub][0][x][2];
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < PER; t++)
        if (rk[t] >= 0) {
            forces[sub][0][rk[t]][0] = w[t][0];
            forces[sub][0][rk[t]][1] = w[t][1];
            forces[sub][0][rk[t]][2] = w[t][2];
        }
    __syncthreads();
    if (!live || c != 0) return;
    const int mk = *S.mark;
    bool tau_on = false;
    if (mk >= 0) {
        tau_on = vnorm(S.fpi[0][a], S.fpi[1][a], S.fpi[2][a]) != 0.0 &&
                 vnorm(S.fpi[0][mk], S.fpi[1][mk], S.fpi[2][mk]) != 0.0;
    }
    double fo[3], tau = 0.0;
    if (!fits) {
        atomicOr(overflow, 1);
        contact_slow(P, S, start, ids, a, o, pp, q, cb, rmp, tau_on, fo, &tau);
    } else {
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int x = 0; x < tl; x++) {
            const double *f = forces[sub][0][x];
            fx = fx + f[0];
            fy = fy + f[1];
            fz = fz + f[2];
        }
        for (int x = 0; x < tu; x++) {
            const double *f = forces[sub][1][x];
            fx = fx + f[0];
            fy = fy + f[1];
            fz = fz + f[2];
            if (tau_on) tau += f[3];
        }
        fo[0] = fx;
        fo[1] = fy;
        fo[2] = fz;
    }
    S.fpp[0][a] = fo[0];
    S.fpp[1][a] = fo[1];
    S.fpp[2][a] = fo[2];
    S.tau_i[a] = tau;
}

// Launch of k_contact_par for n particles (the lane count by n).
inline void launch_contact(hipStream_t st, int n, const PartParams &P, const PartState &S, const int *start,
                           const int *ids, double rmp, const int *t_dev, double ramp_t, int *overflow)
{
    const unsigned ppb = n >= CP_WIDE_MIN ? ContactShape<8>::PPB : ContactShape<32>::PPB;
    const dim3 grid((unsigned)(n <= 0 ? 1 : (n + ppb - 1) / ppb));
    if (n >= CP_WIDE_MIN)
        hipLaunchKernelGGL(k_contact_par<8>, grid, dim3(256), 0, st, P, S, start, ids, rmp, t_dev, ramp_t, overflow);
    else
        hipLaunchKernelGGL(k_contact_par<32>, grid, dim3(256), 0, st, P, S, start, ids, rmp, t_dev, ramp_t, overflow);
}

// Contact pairs (a < b) per particle: count pass then write pass.
__global__ void k_pairs(PartParams P, PartState S, const int *__restrict__ start, const int *__restrict__ ids,
                        const int *__restrict__ pstart, int *__restrict__ pcount, int *__restrict__ pairs,
                        int64_t max_pairs)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.n) return;
    if (S.cell[a] < 0) {  // outside the sorting range: flagged, no pairs
        if (!pstart) pcount[a] = 0;
        return;
    }
    int o, pp, q;
    cell_of(P, S.p[0][a], S.p[1][a], S.p[2][a], &o, &pp, &q);
    const int cb = part_box(S.box, a) * P.ncell_box;
    int cnt = 0;
    const int base = pstart ? pstart[a] : 0;
    for (int k = q - 1; k <= q + 1; k++)
        for (int j = pp - 1; j <= pp + 1; j++)
            for (int i = o - 1; i <= o + 1; i++) {
                const int co = cb + wrapc(i, P.ncx) + wrapc(j, P.ncy) * P.ncx + wrapc(k, P.ncz) * P.ncx * P.ncy;
                for (int e = start[co]; e < start[co + 1]; e++) {
                    const int b = ids[e];
                    if (b <= a) continue;
                    double d[3];
                    pair_disp(P, S, a, b, d);
                    if (vnorm(d[0], d[1], d[2]) < P.r_eq) {
                        if (pstart && (int64_t)(base + cnt) < max_pairs) {
                            pairs[2 * (base + cnt)] = a;
                            pairs[2 * (base + cnt) + 1] = b;
                        }
                        cnt++;
                    }
                }
            }
    if (!pstart) pcount[a] = cnt;
}

// Sort each particle's pair segment by partner id.
__global__ void k_pairs_sort(int n, const int *__restrict__ pstart, int *__restrict__ pairs, int64_t max_pairs)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int b = pstart[a];
    const int e = (int)min((int64_t)pstart[a + 1], max_pairs);
    for (int i = b + 1; i < e; i++) {
        const int v = pairs[2 * i + 1];
        int j = i - 1;
        while (j >= b && pairs[2 * j + 1] > v) { pairs[2 * (j + 1) + 1] = pairs[2 * j + 1]; j--; }
        pairs[2 * (j + 1) + 1] = v;
    }
}

// ---------------------------------------------------------------------------
// Trilinear interpolation, field::interpolate (field.cpp:223-248)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ double interp1(const PartParams &P, const T *__restrict__ v, double px, double py,
                                          double pz)
{
    const int Nx = P.nx, Ny = P.ny, Nz = P.nz;
    const int i = (int)floor(px / P.dx), j = (int)floor(py / P.dy), k = (int)floor(pz / P.dz);
    const double m = px / P.dx - floor(px / P.dx);
    const double n = py / P.dy - floor(py / P.dy);
    const double o = pz / P.dz - floor(pz / P.dz);
    const int i1 = (i + 1) % Nx, j1 = (j + 1) % Ny, k1 = (k + 1) % Nz;
    const size_t sy = (size_t)Nx, sz = (size_t)Nx * Ny;
    double r;
    r = (1.0 - m) * (1.0 - n) * (1.0 - o) * (double)v[i + j * sy + k * sz];
    r += (1.0 - m) * n * (1.0 - o) * (double)v[i + j1 * sy + k * sz];
    r += m * (1.0 - n) * (1.0 - o) * (double)v[i1 + j * sy + k * sz];
    r += m * n * (1.0 - o) * (double)v[i1 + j1 * sy + k * sz];
    r += (1.0 - m) * (1.0 - n) * o * (double)v[i + j * sy + k1 * sz];
    r += (1.0 - m) * n * o * (double)v[i + j1 * sy + k1 * sz];
    r += m * (1.0 - n) * o * (double)v[i1 + j * sy + k1 * sz];
    r += m * n * o * (double)v[i1 + j1 * sy + k1 * sz];
    return r;
}

// interp1 split in two: the eight weights and corner offsets of a position
// (computed once), then the weighted sum of one field, in interp1's order
// and association; several fields at one position share the first part,
// and the corners are 32-bit byte offsets from the field's base (a box's
// field is below 4 GiB), so the gathers need no 64-bit vector addresses.
struct Tri {
    double w[8];
    uint32_t off[8];
};

template <typename T>
__device__ __forceinline__ Tri tri_of(const PartParams &P, double px, double py, double pz)
{
    const int Nx = P.nx, Ny = P.ny, Nz = P.nz;
    const int i = (int)floor(px / P.dx), j = (int)floor(py / P.dy), k = (int)floor(pz / P.dz);
    const double m = px / P.dx - floor(px / P.dx);
    const double n = py / P.dy - floor(py / P.dy);
    const double o = pz / P.dz - floor(pz / P.dz);
    const int i1 = (i + 1) % Nx, j1 = (j + 1) % Ny, k1 = (k + 1) % Nz;
    const uint32_t sy = (uint32_t)Nx, sz = (uint32_t)Nx * (uint32_t)Ny, es = (uint32_t)sizeof(T);
    Tri t;
    t.w[0] = (1.0 - m) * (1.0 - n) * (1.0 - o);
    t.w[1] = (1.0 - m) * n * (1.0 - o);
    t.w[2] = m * (1.0 - n) * (1.0 - o);
    t.w[3] = m * n * (1.0 - o);
    t.w[4] = (1.0 - m) * (1.0 - n) * o;
    t.w[5] = (1.0 - m) * n * o;
    t.w[6] = m * (1.0 - n) * o;
    t.w[7] = m * n * o;
    t.off[0] = (i + j * sy + k * sz) * es;
    t.off[1] = (i + j1 * sy + k * sz) * es;
    t.off[2] = (i1 + j * sy + k * sz) * es;
    t.off[3] = (i1 + j1 * sy + k * sz) * es;
    t.off[4] = (i + j * sy + k1 * sz) * es;
    t.off[5] = (i + j1 * sy + k1 * sz) * es;
    t.off[6] = (i1 + j * sy + k1 * sz) * es;
    t.off[7] = (i1 + j1 * sy + k1 * sz) * es;
    return t;
}

template <typename T>
__device__ __forceinline__ double tri_sum(const Tri &t, const T *__restrict__ v)
{
    const char *b = reinterpret_cast<const char *>(v);
    double r = t.w[0] * (double)*reinterpret_cast<const T *>(b + t.off[0]);
#pragma unroll
    for (int q = 1; q < 8; q++) r += t.w[q] * (double)*reinterpret_cast<const T *>(b + t.off[q]);
    return r;
}

// the eight corner values of one field, and their weighted sum (tri_sum's
// arithmetic on values already gathered)
template <typename T>
__device__ __forceinline__ void tri_gather(const Tri &t, const T *__restrict__ v, T *out)
{
    const char *b = reinterpret_cast<const char *>(v);
#pragma unroll
    for (int q = 0; q < 8; q++) out[q] = *reinterpret_cast<const T *>(b + t.off[q]);
}

template <typename T>
__device__ __forceinline__ double tri_dot(const Tri &t, const T *v)
{
    double r = t.w[0] * (double)v[0];
#pragma unroll
    for (int q = 1; q < 8; q++) r += t.w[q] * (double)v[q];
    return r;
}

// field::interpolate2, field.cpp:251-281 (size 3): the reference's nodes
// are floor+1 and floor+2 ((1-size)/2 = -1 is subtracted) with the
// fractions scaled to m/3 + 1/3
template <typename T>
__device__ __forceinline__ double interp2(const PartParams &P, const T *__restrict__ v, double px, double py,
                                          double pz)
{
    const int Nx = P.nx, Ny = P.ny, Nz = P.nz;
    const int i = (int)floor(px / P.dx), j = (int)floor(py / P.dy), k = (int)floor(pz / P.dz);
    double m = px / P.dx - floor(px / P.dx);
    double n = py / P.dy - floor(py / P.dy);
    double o = pz / P.dz - floor(pz / P.dz);
    m = m / 3 + 1.0 / 3;
    n = n / 3 + 1.0 / 3;
    o = o / 3 + 1.0 / 3;
    const int i0 = (i + 1 + Nx) % Nx, i1 = (i + 2) % Nx;
    const int j0 = (j + 1 + Ny) % Ny, j1 = (j + 2) % Ny;
    const int k0 = (k + 1 + Nz) % Nz, k1 = (k + 2) % Nz;
    const size_t sy = (size_t)Nx, sz = (size_t)Nx * Ny;
    double r;
    r = (1.0 - m) * (1.0 - n) * (1.0 - o) * (double)v[i0 + j0 * sy + k0 * sz];
    r += (1.0 - m) * n * (1.0 - o) * (double)v[i0 + j1 * sy + k0 * sz];
    r += m * (1.0 - n) * (1.0 - o) * (double)v[i1 + j0 * sy + k0 * sz];
    r += m * n * (1.0 - o) * (double)v[i1 + j1 * sy + k0 * sz];
    r += (1.0 - m) * (1.0 - n) * o * (double)v[i0 + j0 * sy + k1 * sz];
    r += (1.0 - m) * n * o * (double)v[i0 + j1 * sy + k1 * sz];
    r += m * (1.0 - n) * o * (double)v[i1 + j0 * sy + k1 * sz];
    r += m * n * o * (double)v[i1 + j1 * sy + k1 * sz];
    return r;
}

// element::get_f_rand (element.cpp:109-144), one component: g =
// sqrt(-2 log r1) cos(2 pi r2), drawn again while |g| > 5, with r1, r2 in
// {0, .., 999}/1000 from a counter-based draw keyed by (seed, step t, Heun
// stage, particle a, component d) in place of rand() % 1000; g = 0 after 64
// draws (an r1 of 0 is the only rejection, probability 1e-3 a draw)
__device__ __forceinline__ double brownian_g(uint64_t seed, int t, int stage, int a, int d, double pi)
{
    uint64_t key = splitmix64(seed);
    key = splitmix64(key ^ (uint64_t)(uint32_t)t);
    key = splitmix64(key ^ (((uint64_t)(uint32_t)(3 * (stage - 1) + d) << 32) | (uint64_t)(uint32_t)a));
    for (int k = 0; k < 64; k++) {
        const uint64_t u = splitmix64(key + (uint64_t)(k + 1) * 0xD1B54A32D192ED03ull);
        const double r1 = (double)((u >> 32) % 1000u) / 1000;
        const double r2 = (double)((u & 0xffffffffull) % 1000u) / 1000;
        const double g = sqrt(-2 * log(r1)) * cos(2 * pi * r2);
        if (!(fabs(g) > 5.0)) return g;
    }
    return 0.0;
}

__device__ __forceinline__ bool grid_cell_ok(const PartParams &P, double px, double py, double pz)
{
    const int i = (int)floor(px / P.dx), j = (int)floor(py / P.dy), k = (int)floor(pz / P.dz);
    return i >= 0 && i < P.nx && j >= 0 && j < P.ny && k >= 0 && k < P.nz;
}

// ---------------------------------------------------------------------------
// Spreading (cluster.cpp:344-410): grid cell and fractions of a particle;
// the per-line node sums are built in particle_stage.hpp and added by the
// x pass of the force transform.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void spread_corner(const PartParams &P, double px, double py, double pz, int *x, int *y,
                                              int *z, double *o, double *p, double *q)
{
    *x = (int)floor(px / P.dx);
    *y = (int)floor(py / P.dy);
    *z = (int)floor(pz / P.dz);
    *o = px / P.dx - floor(px / P.dx);
    *p = py / P.dy - floor(py / P.dy);
    *q = pz / P.dz - floor(pz / P.dz);
}

// Dense spread (unit entry point): one thread per grid line adds its
// per-node sums into a zeroed 3-component grid.
template <typename T>
__global__ void k_spread_dense(int nlines, int nx, size_t G, const int *__restrict__ line_start,
                               const int *__restrict__ node_cnt, const SpreadNode<T> *__restrict__ nodes,
                               T *__restrict__ grid3)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    const int b = line_start[l], cnt = node_cnt[l];
    for (int k = 0; k < cnt; k++) {
        const SpreadNode<T> nd = nodes[b + k];
        const size_t node = (size_t)l * nx + nd.x;
        grid3[node] += nd.v[0];
        grid3[G + node] += nd.v[1];
        grid3[2 * G + node] += nd.v[2];
    }
}

}  // namespace fipi
