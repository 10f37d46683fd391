// One Heun particle stage of the time step in one launch, and the spread
// of the particle forces in one more:
//
//   k_stage     per tile of contact cells, staged in LDS: it_mark
//               (cluster.cpp:128-138) from the argmax partials of the
//               positions the current cell table holds; per particle the
//               trilinear interpolation of c and grad c (element.cpp:55-71)
//               and f_pi (element.cpp:75-97), the contact forces over the 27
//               neighbour cells (cluster.cpp:174-275), summed in the
//               reference's order, the Heun stage and the periodic wrap
//               (element.cpp:156-186); the new position is appended to the
//               next table (the sort of the following stage,
//               cluster.cpp:116-169); in stage 2 the particle id is also
//               appended to the four grid lines its spread touches.
//               Particles are visited tile by tile, so neighbouring threads
//               gather from neighbouring grid nodes.
//   k_spread_nodes  per grid line: the two spread entries of every particle
//               on the line, sorted by (x, particle id) and summed per node
//               in ascending particle order (cluster.cpp:344-410); the nodes
//               go to a pool the force pass reads.
//
// Tables have a fixed capacity per cell / line; entries beyond it go to an
// overflow list that exact slow paths scan, so any input is handled.  Slot
// order inside a cell or line depends on atomic timing, but every sum is
// taken in an order fixed by particle ids, so results are deterministic and
// bitwise those of the reference restatement.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "particle_kernels.hpp"

namespace fipi {

constexpr int CELL_CAP = 16;  // records per contact cell before overflow
constexpr int LINE_CAP = 32;  // particles per grid line before overflow

// A particle as the cell table holds it: its position at the time of the
// append and its id (next: in the overflow list, the next record of the
// same cell, -1 after the last).
struct __align__(16) CellRec {
    double x, y, z;
    int id, next;
};

// Argmax-y partial of one workgroup's particles (cluster.cpp:128-138: the
// strictly largest y > 0, first index on ties; idx < 0: none) with that
// particle's position, from which the next stage takes it_mark and the
// mark's attachment force.
struct __align__(16) MarkPart {
    double x, y, z;
    int idx, pad;
};

struct CellTab {
    int *cnt;       // [ncell] records appended (may exceed CELL_CAP)
    CellRec *rec;   // [ncell * CELL_CAP]
    int *ovf_cnt;   // records in the overflow list
    CellRec *ovf;   // [n] records beyond a cell's capacity
    int *ovf_head;  // [ncell] each cell's first overflow record (-1: none);
                    // -1 wherever cnt is 0
    MarkPart *mp;   // argmax-y partials of the appended positions, one per
                    // workgroup of the appending launch
};

// The three cell tables of the particle stages.  Stage s of a run (s = 2t +
// stage - 1, t the device step counter) reads table s % 3, appends to
// (s + 1) % 3 and empties (s + 2) % 3, the table the stage before it read.
// The choice is made on the device, so a captured step replays unchanged.
struct TabSet {
    CellTab t[3];
};

// What Heun stage 1 hands to stage 2 for one particle, one contiguous block
// indexed by id (written once by stage 1, read once by stage 2): the 32
// grid values stage 1 gathered (c and grad c at the 8 corners, as T) and
// their corner 0 (byte offset in the box's field), p_old and u_old, the
// position after stage 1, the radius and cos theta.
template <typename T>
struct alignas(128) Carry {
    T nodes[32];
    double po[3], uo[3], p1[3], r, ct;
    uint32_t gnode, pad;
};

// The step's results for one particle, written by stage 2 in one 128-byte
// block: the velocity of stage 2, f_pp, f_pi, grad c and c at the
// particle, and its tau term.  The id-indexed arrays of PartState are
// formed from these blocks and the Carry blocks on demand (k_unpack).
struct alignas(128) StepOut {
    double u[3], fpp[3], fpi[3], dcdx[3], cloc, tau;
};

// argmax-y candidate: (y descending, index ascending) is a total order
// with "none" (i < 0) below every candidate, so combines in any order give
// the same result
struct MarkC {
    double x, y, z;
    int i;
};

__device__ __forceinline__ void mark_take(MarkC &m, const MarkC &o)
{
    if (o.i >= 0 && (m.i < 0 || o.y > m.y || (o.y == m.y && o.i < m.i))) m = o;
}

// combine one candidate per thread over a 256-thread workgroup; every
// thread gets the result (sm: 4 LDS slots, free again on return)
__device__ __forceinline__ MarkC block_mark(MarkC m, MarkC *sm)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const MarkC q{__shfl_xor(m.x, o), __shfl_xor(m.y, o), __shfl_xor(m.z, o), __shfl_xor(m.i, o)};
        mark_take(m, q);
    }
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    MarkC r = sm[0];
#pragma unroll
    for (int w = 1; w < 4; w++) mark_take(r, sm[w]);
    __syncthreads();
    return r;
}

struct LineTab {
    int *cnt;        // [nlines] particles appended (may exceed LINE_CAP)
    void *ent;       // [nlines * LINE_CAP] LineEnt<T>
    int *ovf_cnt;    // entries in the overflow list
    int2 *ovf;       // [4n] (particle, next) beyond a line's capacity: each
                     // line's chain, next = -1 after the last
    int *ovf_head;   // [nlines] first overflow entry of each line (-1: none);
                     // -1 wherever cnt is 0
    int *node_base;  // [nlines] first node of each line in the pool
    int *node_cnt;   // [nlines]
    int *pool_top;   // nodes allocated
};

__device__ __forceinline__ void cell_append(const PartParams &P, CellTab T, int g, double x, double y, double z,
                                            int id)
{
    const int slot = atomicAdd(&T.cnt[g], 1);
    CellRec r;
    r.x = x;
    r.y = y;
    r.z = z;
    r.id = id;
    r.next = -1;
    if (slot < CELL_CAP) {
        T.rec[(size_t)g * CELL_CAP + slot] = r;
    } else {
        // pushed on the cell's chain (chain order is arbitrary: every
        // reader orders records by id)
        const int o = atomicAdd(T.ovf_cnt, 1);
        r.next = atomicExch(&T.ovf_head[g], o);
        T.ovf[o] = r;
    }
}

// the four grid lines of the spread of a particle at (px, py, pz)
// (cluster.cpp:368-406: corners at y or y+1 and z or z+1); false when the
// particle lies outside the grid
__device__ __forceinline__ bool spread_lines(const PartParams &P, double px, double py, double pz, int box,
                                             int *l4)
{
    const int x = (int)floor(px / P.dx), y = (int)floor(py / P.dy), z = (int)floor(pz / P.dz);
    if (x < 0 || x >= P.nx || y < 0 || y >= P.ny || z < 0 || z >= P.nz) return false;
    const int y1 = (y + 1) % P.ny, z1 = (z + 1) % P.nz, lb = box * P.nlines_box;
    l4[0] = lb + y + P.ny * z;
    l4[1] = lb + y1 + P.ny * z;
    l4[2] = lb + y + P.ny * z1;
    l4[3] = lb + y1 + P.ny * z1;
    return true;
}

// The spread entries of particle a on line (yl, zl): the corners of
// cluster.cpp:368-406 whose (y, z) is the line's, at x and x+1, with the
// weights w = f * (1-o) * (1-p) * (1-q) etc. evaluated in that order;
// f = (f_pp + f_ext)/dx/dy/dz, f_ext = 0.
template <typename T>
__device__ __forceinline__ void line_entries(const PartParams &P, double px, double py, double pz, double f0,
                                             double f1, double f2, int yl, int zl, int *x0, int *x1, T *w0, T *w1)
{
    int x, y, z;
    double o, p, q;
    spread_corner(P, px, py, pz, &x, &y, &z, &o, &p, &q);
    const double f[3] = {(f0 + 0.0) / P.dx / P.dy / P.dz, (f1 + 0.0) / P.dx / P.dy / P.dz,
                         (f2 + 0.0) / P.dx / P.dy / P.dz};
    const bool ylo = y == yl, zlo = z == zl;
    *x0 = x;
    *x1 = (x + 1) % P.nx;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double a, b;  // at x, at x1
        if (ylo && zlo) {        // corners 0 (x) and 1 (x1)
            a = f[d] * (1 - o) * (1 - p) * (1 - q);
            b = f[d] * o * (1 - p) * (1 - q);
        } else if (!ylo && zlo) {  // corners 6 (x) and 2 (x1)
            a = f[d] * (1 - o) * p * (1 - q);
            b = f[d] * o * p * (1 - q);
        } else if (ylo) {          // corners 5 (x) and 3 (x1)
            a = f[d] * (1 - o) * (1 - p) * q;
            b = f[d] * o * (1 - p) * q;
        } else {                   // corners 7 (x) and 4 (x1)
            a = f[d] * (1 - o) * p * q;
            b = f[d] * o * p * q;
        }
        w0[d] = (T)a;
        w1[d] = (T)b;
    }
}

// A particle's two spread entries on one grid line, formed when it is
// appended, so that the node kernel reads them in place: x of the first
// node (the second is x + 1, periodic), the particle id, and the weighted
// forces at the two nodes (line_entries).
template <typename T>
struct LineEnt {
    int x0, id;
    T w0[3], w1[3];
};

// Appends particle id at (px, py, pz) with force (f0, f1, f2) to the four
// lines l4 of its spread (overflow: (line, id), the exact path recomputes
// the entries from the particle arrays).
template <typename T>
__device__ __forceinline__ void line_append4(const PartParams &P, LineTab Lt, const int *l4, double px, double py,
                                             double pz, double f0, double f1, double f2, int id)
{
    LineEnt<T> *ent = reinterpret_cast<LineEnt<T> *>(Lt.ent);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int line = l4[k];
        const int slot = atomicAdd(&Lt.cnt[line], 1);
        if (slot < LINE_CAP) {
            const int lb = line % P.nlines_box;
            LineEnt<T> e;
            int x1;
            line_entries<T>(P, px, py, pz, f0, f1, f2, lb % P.ny, lb / P.ny, &e.x0, &x1, e.w0, e.w1);
            e.id = id;
            ent[(size_t)line * LINE_CAP + slot] = e;
        } else {
            const int o = atomicAdd(Lt.ovf_cnt, 1);
            Lt.ovf[o] = make_int2(id, atomicExch(&Lt.ovf_head[line], o));
        }
    }
}

// ---------------------------------------------------------------------------
// Table from positions (after the state is set or restored) and its argmax
// partials, one per workgroup (launched with as many workgroups as the
// stage kernel); the table's counters must be zero.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cells_append(PartParams P, const double *__restrict__ px,
                                                      const double *__restrict__ py, const double *__restrict__ pz,
                                                      const int *__restrict__ box, CellTab T, int *__restrict__ err)
{
    __shared__ MarkC sm[4];
    MarkC m{0.0, 0.0, 0.0, -1};
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < P.n; a += gridDim.x * blockDim.x) {
        const double x = px[a], y = py[a], z = pz[a];
        if (y > 0.0) mark_take(m, MarkC{x, y, z, a});
        int i, j, k;
        const int cc = cell_of(P, x, y, z, &i, &j, &k);
        if (cc < 0) {
            atomicOr(err, 1);
            continue;
        }
        cell_append(P, T, cc + part_box(box, a) * P.ncell_box, x, y, z, a);
    }
    const MarkC r = block_mark(m, sm);
    if (threadIdx.x == 0) T.mp[blockIdx.x] = MarkPart{r.x, r.y, r.z, r.i, 0};
}

// counters of a table (and optionally of the spread lines) to zero, and
// a cell table's overflow chains (head, optional) to empty; a kernel, not a
// memset, so that graph capture records no memset node
__global__ void k_tab_clear(int *cnt, int n, int *ovf_cnt, int *line_ovf, int *pool_top, int *head)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    for (int q = i; q < n; q += gridDim.x * blockDim.x) {
        cnt[q] = 0;
        if (head) head[q] = -1;
    }
    if (i == 0) {
        *ovf_cnt = 0;
        if (line_ovf) *line_ovf = 0;
        if (pool_top) *pool_top = 0;
    }
}

// spread lines of n particles with forces f (the spread outside the step's
// stage 2)
template <typename T>
__global__ __launch_bounds__(256) void k_spread_lines(PartParams P, const double *__restrict__ px,
                                                      const double *__restrict__ py, const double *__restrict__ pz,
                                                      const double *__restrict__ f0, const double *__restrict__ f1,
                                                      const double *__restrict__ f2, const int *__restrict__ box,
                                                      LineTab Lt, int *__restrict__ err)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.n) return;
    int l4[4];
    const double x = px[a], y = py[a], z = pz[a];
    if (!spread_lines(P, x, y, z, part_box(box, a), l4)) {
        atomicOr(err, 2);
        return;
    }
    line_append4<T>(P, Lt, l4, x, y, z, f0[a], f1[a], f2[a], a);
}

// ---------------------------------------------------------------------------
// Shared per-particle arithmetic
// ---------------------------------------------------------------------------
// element::get_f_pi (element.cpp:75-97): the capillary attachment force of
// a sphere of radius r at local phase field phi and gradient (gx, gy, gz);
// ct = cos(contact angle) centres the spring on r ct (0: the reference)
__device__ __forceinline__ void capillary_force(const PartParams &P, double phi, double gx, double gy, double gz,
                                                double r, double ct, double *f)
{
    f[0] = f[1] = f[2] = 0.0;
    if (fabs(phi) <= P.tanh2) {
        const double gn = vnorm(gx, gy, gz);
        if (gn != 0.0) {
            double d = 0.5 * log((1.0 + phi) / (1.0 - phi)) * P.sqrt2 * P.epsilon;
            if (ct != 0.0) d = d - r * ct;
            double fm = 0.0;
            if (fabs(d) < r) fm = -P.a_pi * P.pi * d * P.sigma;
            f[0] = gx / gn * fm;
            f[1] = gy / gn * fm;
            f[2] = gz / gn * fm;
        }
    }
}

// element::advect (element.cpp:156-173), one component of the stage
// velocity uf + (f_pi + f_ext + f_pp)/(6 pi mu r); f_ext and uf are zero in
// the reference step
__device__ __forceinline__ double heun_velocity(const PartParams &P, double fpi, double fext, double fpp, double r,
                                                double uf)
{
    const double fs = fpi + fext + fpp;
    return uf + fs / 6.0 / P.pi / P.mu / r;
}

// ---------------------------------------------------------------------------
// Stage kernel: cell records and contact sums
// ---------------------------------------------------------------------------
// Visit the records of global cell co: the table part in slot order, then
// the cell's overflow chain: f(rec, ref) with ref >= 0 a table slot and
// ref < 0 overflow entry -1-ref.
template <class F>
__device__ __forceinline__ void cell_visit(const CellTab &T, int co, F &&f)
{
    const int m = T.cnt[co];
    const int mt = m < CELL_CAP ? m : CELL_CAP;
    for (int e = 0; e < mt; e++) f(T.rec[(size_t)co * CELL_CAP + e], co * CELL_CAP + e);
    if (m > CELL_CAP) {
        for (int k = T.ovf_head[co]; k >= 0;) {
            const CellRec r = T.ovf[k];
            f(r, -1 - k);
            k = r.next;
        }
    }
}

__device__ __forceinline__ CellRec cell_rec(const CellTab &T, int ref)
{
    return ref >= 0 ? T.rec[ref] : T.ovf[-1 - ref];
}

// Force on a (at ax, ay, az) from the partner record r; w[3] = 0.5 F.d.
__device__ __forceinline__ void contact_rec(const PartParams &P, double ax, double ay, double az, int a,
                                            const CellRec &r, double rmp, double *w)
{
    double d[3];
    pair_disp_pos(P, ax, ay, az, r.x, r.y, r.z, a, r.id, d);
    const double nrm = vnorm(d[0], d[1], d[2]);
    const double sc = (P.r_eq - nrm) * rmp * P.k_n;
    const double f0 = (d[0] / nrm) * sc, f1 = (d[1] / nrm) * sc, f2 = (d[2] / nrm) * sc;
    w[0] = f0;
    w[1] = f1;
    w[2] = f2;
    w[3] = 0.5 * ((f0 * d[0] + f1 * d[1]) + f2 * d[2]);
}

// Exact reference-order sum over the table (tiles whose records overflow
// the fast path): the partners b < a by repeated scans of the 27 cells for
// the next id, the partners b > a cell by cell.
constexpr int SLOW_NC = 8;  // partners in contact collected per cell in one pass
__device__ void contact_slow_tab(const PartParams &P, const CellTab &T, int a, double ax, double ay, double az,
                                 int o, int pp, int q, int cb, double rmp, bool tau_on, double *fo, double *tau_out)
{
    double fx = 0.0, fy = 0.0, fz = 0.0;
    int last = -1;
    for (;;) {
        int best = 0x7fffffff;
        CellRec br{};
        for (int k = q - 1; k <= q + 1; k++)
            for (int j = pp - 1; j <= pp + 1; j++)
                for (int i = o - 1; i <= o + 1; i++) {
                    const int co = cb + wrapc(i, P.ncx) + wrapc(j, P.ncy) * P.ncx + wrapc(k, P.ncz) * P.ncx * P.ncy;
                    cell_visit(T, co, [&](const CellRec &r, int) {
                        const int b = r.id;
                        if (b >= a || b <= last || b >= best) return;
                        double d[3];
                        pair_disp_pos(P, ax, ay, az, r.x, r.y, r.z, a, b, d);
                        if (vnorm(d[0], d[1], d[2]) < P.r_eq) {
                            best = b;
                            br = r;
                        }
                    });
                }
        if (best == 0x7fffffff) break;
        double w[4];
        contact_rec(P, ax, ay, az, a, br, rmp, w);
        fx = fx + w[0];
        fy = fy + w[1];
        fz = fz + w[2];
        last = best;
    }
    double tau = 0.0;
    for (int k = q - 1; k <= q + 1; k++)
        for (int j = pp - 1; j <= pp + 1; j++)
            for (int i = o - 1; i <= o + 1; i++) {
                const int co = cb + wrapc(i, P.ncx) + wrapc(j, P.ncy) * P.ncx + wrapc(k, P.ncz) * P.ncx * P.ncy;
                // partners b > a of one cell in ascending id (list order):
                // one pass collects those in contact, in id order, when
                // there are at most SLOW_NC; repeated scans otherwise
                int cid[SLOW_NC], cref[SLOW_NC];
                int nc = 0;
                bool more = false;
                cell_visit(T, co, [&](const CellRec &r, int ref) {
                    if (r.id <= a || more) return;
                    double d[3];
                    pair_disp_pos(P, ax, ay, az, r.x, r.y, r.z, a, r.id, d);
                    if (!(vnorm(d[0], d[1], d[2]) < P.r_eq)) return;
                    if (nc == SLOW_NC) {
                        more = true;
                        return;
                    }
                    int s = nc++;
                    for (; s > 0 && cid[s - 1] > r.id; s--) {
                        cid[s] = cid[s - 1];
                        cref[s] = cref[s - 1];
                    }
                    cid[s] = r.id;
                    cref[s] = ref;
                });
                if (!more) {
                    for (int s = 0; s < nc; s++) {
                        double w[4];
                        contact_rec(P, ax, ay, az, a, cell_rec(T, cref[s]), rmp, w);
                        fx = fx + w[0];
                        fy = fy + w[1];
                        fz = fz + w[2];
                        if (tau_on) tau += w[3];
                    }
                    continue;
                }
                int lastc = a;
                for (;;) {
                    int nb = 0x7fffffff;
                    CellRec nr{};
                    cell_visit(T, co, [&](const CellRec &r, int) {
                        if (r.id > lastc && r.id < nb) {
                            nb = r.id;
                            nr = r;
                        }
                    });
                    if (nb == 0x7fffffff) break;
                    lastc = nb;
                    double d[3];
                    pair_disp_pos(P, ax, ay, az, nr.x, nr.y, nr.z, a, nb, d);
                    if (vnorm(d[0], d[1], d[2]) < P.r_eq) {
                        double w[4];
                        contact_rec(P, ax, ay, az, a, nr, rmp, w);
                        fx = fx + w[0];
                        fy = fy + w[1];
                        fz = fz + w[2];
                        if (tau_on) tau += w[3];
                    }
                }
            }
    fo[0] = fx;
    fo[1] = fy;
    fo[2] = fz;
    *tau_out = tau;
}

// ---------------------------------------------------------------------------
// Stage kernel: a tile of contact cells per workgroup
// ---------------------------------------------------------------------------
// A workgroup owns a tile of TILE_X x TILE_Y x TILE_Z contact cells (one
// thread per cell) and stages the records of the tile and its one-cell halo
// (periodic) in LDS, each cell's records sorted by id.  Every thread then
// handles the particles of its cell from LDS: the interpolation and f_pi at
// the particle, a bit per (neighbour cell, slot) marking the partners in
// contact, the partners b < a summed in ascending id and the partners b > a
// in traversal order (neighbour cells k, j, i; ids ascending in a cell), as
// cluster.cpp:174-275 accumulates; then the Heun stage, the wrap, the next
// sort and (stage 2) the spread lines.  A tile whose halo holds an
// overflowing cell or more records than the LDS takes the exact
// global-memory path (contact_slow_tab).
constexpr int TILE_X = 8, TILE_Y = 8, TILE_Z = 2;
constexpr int TILE_CELLS = TILE_X * TILE_Y * TILE_Z;   // 128 cells, ~170 spheres at 20 %
constexpr int HALO_X = TILE_X + 2, HALO_Y = TILE_Y + 2, HALO_Z = TILE_Z + 2;
constexpr int HALO_N = HALO_X * HALO_Y * HALO_Z;       // 400 cells
constexpr int HALO_PER = (HALO_N + 255) / 256;         // halo cells per thread
constexpr int TILE_RECS = 1024;                         // records staged in LDS
constexpr int HIT_WORDS = (27 * CELL_CAP + 63) / 64;    // bits per (neighbour, slot)
static_assert(CELL_CAP == 16, "a neighbour's hit bits stay inside one 64-bit word");
static_assert(TILE_CELLS <= 256, "a thread per tile cell");

// tiles per box along each axis, and the launch grid of a slab
__host__ __device__ inline int tiles_of(int nc, int t) { return (nc + t - 1) / t; }

// (d2, d): displacement of b from a with the reference's minimum image and
// the coincident-pair jitter; in contact when sqrt(d2) < r_eq.  The squared
// distance is the one vnorm() takes the root of, so the test below is the
// reference's, with the root skipped where d2 is clearly beyond r_eq^2.
__device__ __forceinline__ bool in_contact(const PartParams &P, double ax, double ay, double az, int a,
                                           const CellRec &r, double r_eq2_hi)
{
    double d[3] = {ax - r.x, ay - r.y, az - r.z};
    const double L[3] = {P.lx, P.ly, P.lz};
    const double th[3] = {P.dcx * (P.ncx - 2), P.dcy * (P.ncy - 2), P.dcz * (P.ncz - 2)};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        if (d[q] > th[q]) d[q] -= L[q];
        else if (d[q] < -th[q]) d[q] += L[q];
    }
    double d2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
    if (d2 == 0.0) {
        // vnorm(d) == 0: the jitter of the coincident pair (as pair_disp_pos)
        const int b = r.id;
        const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double jv = jitter(lo, hi, q, P.r_p);
            d[q] = a < b ? jv : -jv;
        }
        d2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
    }
    if (d2 > r_eq2_hi) return false;
    return sqrt(d2) < P.r_eq;
}

// LDS index of neighbour n (0..26, traversal order) of halo cell hc, and
// the first and end record of that cell
__device__ __forceinline__ int halo_nb(int hc, int n)
{
    const int dz = n / 9 - 1, dy = (n / 3) % 3 - 1, dx = n % 3 - 1;
    return hc + dx + HALO_X * (dy + HALO_Y * dz);
}

// Grid values of c, grad c (and u) at a particle: field.cpp:223-248 as
// element.cpp:55-71 calls it, box offset applied by the caller.
template <typename T>
struct StageFields {
    const T *c, *g0, *g1, *g2, *u0, *u1, *u2;
};

// it_mark of the sorts of stage 1 and stage 2 (cluster.cpp:128-138: the
// strictly largest y > 0, first index on ties; kept when no y > 0), from the
// argmax partials of the table stage 1 read (positions at the step's start)
// and of the table it filled (positions after stage 1, also in Carry), and
// whether the attachment force of the stage-2 mark is non-zero at its
// position (the tau gate, cluster.cpp:256-259; diagnostics apply it to the
// sum of the particles' tau terms).  One workgroup of stage 2 runs it beside
// its tile (nothing it reads is written during stage 2); every thread of
// that workgroup calls it.
template <typename T>
__device__ void stage_mark(const PartParams &P, const PartState &S, const StageFields<T> &F, const TabSet &tabs,
                           int nparts, int tnow, MarkC *smk)
{
    const int k3 = (2 * tnow) % 3;
    const MarkPart *m0 = tabs.t[k3].mp, *m1 = tabs.t[(k3 + 1) % 3].mp;
    MarkC a{0.0, 0.0, 0.0, -1}, b{0.0, 0.0, 0.0, -1};
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
        const MarkPart q0 = m0[k], q1 = m1[k];
        mark_take(a, MarkC{q0.x, q0.y, q0.z, q0.idx});
        mark_take(b, MarkC{q1.x, q1.y, q1.z, q1.idx});
    }
    a = block_mark(a, smk);
    b = block_mark(b, smk);
    if (threadIdx.x != 0) return;
    int mk = b.i;
    double mp[3] = {b.x, b.y, b.z};
    if (mk < 0) {
        mk = a.i >= 0 ? a.i : *S.mark;
        if (mk >= 0) {
#pragma unroll
            for (int d = 0; d < 3; d++) mp[d] = reinterpret_cast<const Carry<T> *>(S.carry)[mk].p1[d];
        }
    }
    bool on = false;
    if (mk >= 0) {
        *S.mark = mk;
        double phi = 0.0, gr[3] = {0.0, 0.0, 0.0}, f[3];
        if (grid_cell_ok(P, mp[0], mp[1], mp[2])) {
            const size_t bo = (size_t)part_box(S.box, mk) * P.rs;
            const Tri tr = tri_of<T>(P, mp[0], mp[1], mp[2]);
            phi = tri_sum(tr, F.c + bo);
            gr[0] = tri_sum(tr, F.g0 + bo);
            gr[1] = tri_sum(tr, F.g1 + bo);
            gr[2] = tri_sum(tr, F.g2 + bo);
        }
        capillary_force(P, phi, gr[0], gr[1], gr[2], S.rad[mk], P.has_ctheta ? S.cth[mk] : 0.0, f);
        on = vnorm(f[0], f[1], f[2]) != 0.0;
    }
    *S.mark_on = on ? 1 : 0;
}

// The reference's unused physics of one stage, per particle at its
// stage-start position (p for stage 1, Carry::p1 for stage 2), before k_stage:
// uf from u (element.cpp:43-49 / field.cpp:251-281) and the Brownian force
// of this step and stage (element.cpp:109-144).
template <typename T>
__global__ __launch_bounds__(256) void k_fluid_noise(PartParams P, PartState S, StageFields<T> F,
                                                     const int *__restrict__ t_dev, int stage)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.n) return;
    const Carry<T> &cy = reinterpret_cast<const Carry<T> *>(S.carry)[a];
    const double px = stage == 1 ? S.p[0][a] : cy.p1[0], py = stage == 1 ? S.p[1][a] : cy.p1[1],
                 pz = stage == 1 ? S.p[2][a] : cy.p1[2];
    if (P.interp_u && grid_cell_ok(P, px, py, pz)) {
        const size_t bo = (size_t)part_box(S.box, a) * P.rs;
        const T *us[3] = {F.u0 + bo, F.u1 + bo, F.u2 + bo};
#pragma unroll
        for (int d = 0; d < 3; d++)
            S.uf[d][a] = P.interp_u == 1 ? interp1(P, us[d], px, py, pz) : interp2(P, us[d], px, py, pz);
    }
    if (P.brown_amp > 0.0) {
        const int t = *t_dev;
#pragma unroll
        for (int d = 0; d < 3; d++) S.frand[d][a] = P.brown_amp * brownian_g(P.noise_seed, t, stage, a, d, P.pi);
    }
}

// lines_on: stage 1 resets the spread's overflow list and node pool for
// stage 2; stage 2 appends each particle to its spread lines.  Three waves
// per SIMD (168 VGPRs): the tiles of a 256^3 box then fit the GPU at once.
template <int STAGE, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_stage(PartParams P, PartState S, StageFields<T> F, TabSet tabs, int ncell,
                                               LineTab lt, int lines_on, const int *__restrict__ t_dev,
                                               double ramp_t, int *__restrict__ overflow)
{
    __shared__ int hoff[HALO_N + 1];
    // the halo's records: exact positions, and an FP32 copy with the id (w)
    // for the reject test of the contact search
    __shared__ double rxs[TILE_RECS], rys[TILE_RECS], rzs[TILE_RECS];
    __shared__ float4 rq[TILE_RECS];
    __shared__ int plist[TILE_RECS];   // the tile's particles: LDS record
    __shared__ short pcell[TILE_RECS];  // and halo cell
    __shared__ int wsum[4];
    __shared__ MarkC smk[4];
    auto rec_at = [&](int e) { return CellRec{rxs[e], rys[e], rzs[e], __float_as_int(rq[e].w), 0}; };
    const int tid = threadIdx.x;
    const int ntx = tiles_of(P.ncx, TILE_X), nty = tiles_of(P.ncy, TILE_Y), ntz = tiles_of(P.ncz, TILE_Z);
    const int box = blockIdx.z / ntz;
    const int x0 = blockIdx.x * TILE_X, y0 = blockIdx.y * TILE_Y, z0 = (blockIdx.z % ntz) * TILE_Z;
    const int cb = box * P.ncell_box;
    const int tnow = *t_dev;
    const double rmp = ramp_dev((double)(tnow + 1), ramp_t);
    // slightly above r_eq^2: below it the exact root decides
    const double r_eq2_hi = P.r_eq * P.r_eq * (1.0 + 1e-12);
    (void)nty;
    (void)ntx;
    const int nwg = (int)(gridDim.x * gridDim.y * gridDim.z);
    const int wg = (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
    const int k3 = (2 * tnow + STAGE - 1) % 3;
    const CellTab cur = tabs.t[k3], nxt = tabs.t[(k3 + 1) % 3], clr = tabs.t[(k3 + 2) % 3];
    // 0. the table the next stage appends to, emptied (the stage before this
    // one read it); stage 1 also resets the spread lines' shared lists
    for (int q = wg * 256 + tid; q < ncell; q += nwg * 256) {
        clr.cnt[q] = 0;
        clr.ovf_head[q] = -1;
    }
    if (wg == 0 && tid == 0) {
        *clr.ovf_cnt = 0;
        if (STAGE == 1 && lines_on) {
            *lt.ovf_cnt = 0;
            *lt.pool_top = 0;
        }
    }
    const size_t bo = (size_t)box * P.rs;
    // stage 2: it_mark and the tau gate, in the first workgroup
    if (STAGE == 2 && wg == 0) stage_mark<T>(P, S, F, tabs, nwg, tnow, smk);
    // 1. record counts of the halo cells (HALO_PER consecutive cells a thread)
    int cntv[HALO_PER], cov[HALO_PER];
    int over = 0;
#pragma unroll
    for (int q = 0; q < HALO_PER; q++) {
        const int h = HALO_PER * tid + q;
        cntv[q] = 0;
        cov[q] = 0;
        if (h < HALO_N) {
            const int hx = h % HALO_X, hy = (h / HALO_X) % HALO_Y, hz = h / (HALO_X * HALO_Y);
            const int gx = wrapc(x0 + hx - 1, P.ncx), gy = wrapc(y0 + hy - 1, P.ncy), gz = wrapc(z0 + hz - 1, P.ncz);
            cov[q] = cb + gx + P.ncx * (gy + P.ncy * gz);
            const int m = cur.cnt[cov[q]];
            cntv[q] = m < CELL_CAP ? m : CELL_CAP;
            if (m > CELL_CAP) over = 1;
        }
    }
    const bool any_over = __syncthreads_or(over) != 0;
    // 2. exclusive offsets over the halo (block scan of per-thread sums)
    int mysum = 0;
#pragma unroll
    for (int q = 0; q < HALO_PER; q++) mysum += cntv[q];
    const int lane = tid & 63, wid = tid >> 6;
    const int inc = wave_incl_scan(mysum);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int off = inc - mysum;
    for (int w = 0; w < wid; w++) off += wsum[w];
    const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
#pragma unroll
    for (int q = 0; q < HALO_PER; q++) {
        const int h = HALO_PER * tid + q;
        if (h < HALO_N) hoff[h] = off;
        off += cntv[q];
    }
    if (tid == 0) hoff[HALO_N] = total;
    const bool fb = any_over || total > TILE_RECS;
    // 3. records into LDS, each cell's sorted by id (a cell holds a few)
    __syncthreads();
    if (!fb) {
#pragma unroll
        for (int q = 0; q < HALO_PER; q++) {
            const int h = HALO_PER * tid + q;
            if (h >= HALO_N) continue;
            const int b0 = hoff[h], m = cntv[q];
            // insertion by id while loading (a cell holds a few records)
            for (int e = 0; e < m; e++) {
                const CellRec v = cur.rec[(size_t)cov[q] * CELL_CAP + e];
                int j = b0 + e - 1;
                while (j >= b0 && __float_as_int(rq[j].w) > v.id) {
                    rxs[j + 1] = rxs[j];
                    rys[j + 1] = rys[j];
                    rzs[j + 1] = rzs[j];
                    rq[j + 1] = rq[j];
                    j--;
                }
                rxs[j + 1] = v.x;
                rys[j + 1] = v.y;
                rzs[j + 1] = v.z;
                rq[j + 1] = make_float4((float)v.x, (float)v.y, (float)v.z, __int_as_float(v.id));
            }
        }
    }
    __syncthreads();
    // 4. the tile's particles, one per thread: a block scan over the tile
    // cells' counts lists them (LDS record and halo cell of each)
    const int lx = tid % TILE_X, ly = (tid / TILE_X) % TILE_Y, lz = tid / (TILE_X * TILE_Y);
    const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
    const bool cell_on = tid < TILE_CELLS && gx < P.ncx && gy < P.ncy && gz < P.ncz;
    const int hc_t = (lx + 1) + HALO_X * ((ly + 1) + HALO_Y * (lz + 1));
    const int co = cb + gx + P.ncx * (gy + P.ncy * gz);
    MarkC mine{0.0, 0.0, 0.0, -1};  // argmax-y of this thread's particles after the stage
    // Per particle before the contact search: the interpolation of c and
    // grad c at its stage-start position and f_pi (interpolate_local_variables
    // and compute_f_pi, main.cpp:198-200), and the inputs of the stage's end.
    // Stage 2 stores the particle's local values (the step's last ones).  uf
    // and the Brownian force (off in the reference step) come from
    // k_fluid_noise, launched before this kernel when they are on.
    struct Pre {
        double fpi[3], r, ct, po[3], uo[3];
    };
    const bool with_uf = P.interp_u != 0, with_noise = P.brown_amp > 0.0;
    Carry<T> *const carry = reinterpret_cast<Carry<T> *>(S.carry);
    StepOut *const sout = reinterpret_cast<StepOut *>(S.out);
    auto prefetch = [&](int a, double ax, double ay, double az) {
        Pre q;
        const Carry<T> &cy = carry[a];
        if (STAGE == 1) {
            q.r = S.rad[a];
            q.ct = P.has_ctheta ? S.cth[a] : 0.0;
        } else {
            q.r = cy.r;
            q.ct = cy.ct;
        }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            q.po[d] = STAGE == 2 ? cy.po[d] : 0.0;
            q.uo[d] = STAGE == 2 ? cy.uo[d] : 0.0;
        }
        // c and grad c at the particle.  Both stages read the same fields
        // (c^{n+1}): stage 1 keeps the 32 corner values it gathers, and stage
        // 2 takes them when the particle is still in the same grid cell (the
        // same corners, so the same values and the same sums), two
        // contiguous lines instead of sixteen scattered ones.
        double phi = 0.0, gr[3] = {0.0, 0.0, 0.0};
        if (!grid_cell_ok(P, ax, ay, az)) {
            atomicOr(S.err, 2);
        } else {
            const Tri tr = tri_of<T>(P, ax, ay, az);
            const T *fld[4] = {F.c + bo, F.g0 + bo, F.g1 + bo, F.g2 + bo};
            double val[4];
            if (STAGE == 2 && cy.gnode == tr.off[0]) {
#pragma unroll
                for (int f = 0; f < 4; f++) val[f] = tri_dot(tr, cy.nodes + 8 * f);
            } else if (STAGE == 1) {
                Carry<T> &cw = carry[a];
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    T v[8];
                    tri_gather(tr, fld[f], v);
#pragma unroll
                    for (int c8 = 0; c8 < 8; c8++) cw.nodes[8 * f + c8] = v[c8];
                    val[f] = tri_dot(tr, v);
                }
                cw.gnode = tr.off[0];
            } else {
#pragma unroll
                for (int f = 0; f < 4; f++) val[f] = tri_sum(tr, fld[f]);
            }
            phi = val[0];
            gr[0] = val[1];
            gr[1] = val[2];
            gr[2] = val[3];
        }
        capillary_force(P, phi, gr[0], gr[1], gr[2], q.r, q.ct, q.fpi);
        if (STAGE == 2) {
            StepOut &o = sout[a];
            o.cloc = phi;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                o.dcdx[d] = gr[d];
                o.fpi[d] = q.fpi[d];
            }
        }
        return q;
    };
    // Stage 1 hands its results to stage 2 in the particle's Carry block;
    // stage 2 writes the step's results to its StepOut block and the new
    // position to p.
    auto finish = [&](int a, double ax, double ay, double az, const double *fo, double tau, const Pre &q) {
        if (STAGE == 2) {
            StepOut &o = sout[a];
            o.fpp[0] = fo[0];
            o.fpp[1] = fo[1];
            o.fpp[2] = fo[2];
            o.tau = tau;
        }
        // Heun stage (element.cpp:156-173), periodic wrap
        // (element.cpp:177-186)
        const double r = q.r;
        const double L[3] = {P.lx, P.ly, P.lz};
        const double p0[3] = {ax, ay, az};
        double np[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double uf = with_uf ? S.uf[d][a] : 0.0;
            const double fext = with_noise ? S.frand[d][a] : 0.0;
            const double v = heun_velocity(P, q.fpi[d], fext, fo[d], r, uf);
            double x;
            if (STAGE == 1) {
                carry[a].uo[d] = v;
                carry[a].po[d] = p0[d];
                x = p0[d] + v * P.dt;
            } else {
                sout[a].u[d] = v;
                x = q.po[d] + (q.uo[d] + v) * (0.5 * P.dt);
            }
            np[d] = fmod(x + L[d], L[d]);
            if (STAGE == 1) carry[a].p1[d] = np[d];
            else S.p[d][a] = np[d];
        }
        if (STAGE == 1) {
            carry[a].r = r;
            carry[a].ct = q.ct;
        }
        // the sort of the next stage (cluster.cpp:116-169)
        int ci, cj, ck;
        const int cc = cell_of(P, np[0], np[1], np[2], &ci, &cj, &ck);
        if (cc < 0) atomicOr(S.err, 1);
        else cell_append(P, nxt, cc + cb, np[0], np[1], np[2], a);
        if (np[1] > 0.0) mark_take(mine, MarkC{np[0], np[1], np[2], a});
        if (STAGE == 2 && lines_on) {
            // the spread of f_pp at the new position for the next step's force
            int l4[4];
            if (!spread_lines(P, np[0], np[1], np[2], box, l4)) atomicOr(S.err, 2);
            else line_append4<T>(P, lt, l4, np[0], np[1], np[2], fo[0], fo[1], fo[2], a);
        }
    };
    // tau counts when f_pi of the particle and of it_mark are non-zero
    // (cluster.cpp:256-259): the particle's part here, it_mark's part
    // (mark_on) where the terms are summed; stage 1 keeps no tau (stage 2
    // overwrites it)
    auto tau_gate = [&](const Pre &q) { return STAGE == 2 && vnorm(q.fpi[0], q.fpi[1], q.fpi[2]) != 0.0; };
    if (fb) {
        // exact path from global memory (overflowing cells or a crowded
        // halo), one tile cell a thread
        if (cell_on) {
            atomicOr(overflow, 1);
            int o_, pp_, q_;
            cell_visit(cur, co, [&](const CellRec &ra, int) {
                const int a = ra.id;
                const Pre q = prefetch(a, ra.x, ra.y, ra.z);
                cell_of(P, ra.x, ra.y, ra.z, &o_, &pp_, &q_);
                double fo[3], tau = 0.0;
                contact_slow_tab(P, cur, a, ra.x, ra.y, ra.z, o_, pp_, q_, cb, rmp, tau_gate(q), fo, &tau);
                finish(a, ra.x, ra.y, ra.z, fo, tau, q);
            });
        }
    } else {
        const int mine = cell_on ? hoff[hc_t + 1] - hoff[hc_t] : 0;
        const int pinc = wave_incl_scan(mine);
        if (lane == 63) wsum[wid] = pinc;
        __syncthreads();
        int pb = pinc - mine;
        for (int w = 0; w < wid; w++) pb += wsum[w];
        const int ptot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        for (int k = 0; k < mine; k++) {
            plist[pb + k] = hoff[hc_t] + k;
            pcell[pb + k] = (short)hc_t;
        }
        __syncthreads();
        // FP32 reject test: a pair within r_eq has |d| < r_eq + 1e-4 in FP32
        // (positions below 2^24 ulps of the box), so only pairs that pass
        // (d^2 <= (r_eq + 1e-3)^2) take the exact FP64 test; where the
        // minimum-image threshold comes within 1e-2 of r_eq every pair takes it
        const float Lf[3] = {(float)P.lx, (float)P.ly, (float)P.lz};
        const float thf[3] = {(float)(P.dcx * (P.ncx - 2)), (float)(P.dcy * (P.ncy - 2)), (float)(P.dcz * (P.ncz - 2))};
        const double thmin = fmin(P.dcx * (P.ncx - 2), fmin(P.dcy * (P.ncy - 2), P.dcz * (P.ncz - 2)));
        const float lim2 = thmin - P.r_eq > 1e-2 ? (float)((P.r_eq + 1e-3) * (P.r_eq + 1e-3)) : 3.0e38f;
        for (int pi = tid; pi < ptot; pi += 256) {
            const CellRec ra = rec_at(plist[pi]);
            const int hc = pcell[pi];
            const int a = ra.id;
            const double ax = ra.x, ay = ra.y, az = ra.z;
            const Pre q = prefetch(a, ax, ay, az);
            const float axf = (float)ax, ayf = (float)ay, azf = (float)az;
            uint64_t hit[HIT_WORDS];
#pragma unroll
            for (int w = 0; w < HIT_WORDS; w++) hit[w] = 0;
            // 9 rows of three x-neighbouring cells, contiguous in LDS
#pragma unroll
            for (int row = 0; row < 9; row++) {
                const int hm = hc + HALO_X * (row % 3 - 1 + HALO_Y * (row / 3 - 1));  // the dx = 0 cell
                const int r0 = hoff[hm - 1], r1 = hoff[hm], r2 = hoff[hm + 1], r3 = hoff[hm + 2];
                for (int e = r0; e < r3; e++) {
                    const float4 bq = rq[e];
                    float dv[3] = {axf - bq.x, ayf - bq.y, azf - bq.z};
#pragma unroll
                    for (int c3 = 0; c3 < 3; c3++)
                        dv[c3] = dv[c3] > thf[c3] ? dv[c3] - Lf[c3] : (dv[c3] < -thf[c3] ? dv[c3] + Lf[c3] : dv[c3]);
                    const float d2f = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
                    if (d2f > lim2) continue;
                    const CellRec rb = rec_at(e);
                    if (rb.id == a) continue;
                    if (in_contact(P, ax, ay, az, a, rb, r_eq2_hi)) {
                        // neighbour n = 3 row + j (j: which of the three cells)
                        const int j = (e >= r1) + (e >= r2);
                        const int k = e - (j == 0 ? r0 : (j == 1 ? r1 : r2));
#pragma unroll
                        for (int jj = 0; jj < 3; jj++)
                            if (j == jj) {
                                const int n = 3 * row + jj;
                                hit[n / 4] |= 1ull << ((n % 4) * CELL_CAP + k);
                            }
                    }
                }
            }
            auto rec_of = [&](int bit) {
                const int n = bit / CELL_CAP;
                return rec_at(hoff[halo_nb(hc, n)] + bit % CELL_CAP);
            };
            double fx = 0.0, fy = 0.0, fz = 0.0, tau = 0.0;
            // partners b < a in ascending id
            int last = -1;
            for (;;) {
                int best = 0x7fffffff, bbit = -1;
#pragma unroll
                for (int w = 0; w < HIT_WORDS; w++) {
                    uint64_t m = hit[w];
                    while (m) {
                        const int bit = w * 64 + __ffsll((unsigned long long)m) - 1;
                        m &= m - 1;
                        const int b = rec_of(bit).id;
                        if (b < a && b > last && b < best) {
                            best = b;
                            bbit = bit;
                        }
                    }
                }
                if (bbit < 0) break;
                double wv[4];
                contact_rec(P, ax, ay, az, a, rec_of(bbit), rmp, wv);
                fx = fx + wv[0];
                fy = fy + wv[1];
                fz = fz + wv[2];
                last = best;
            }
            // partners b > a in traversal order
            const bool tau_on = tau_gate(q);
#pragma unroll
            for (int w = 0; w < HIT_WORDS; w++) {
                uint64_t m = hit[w];
                while (m) {
                    const int bit = w * 64 + __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    const CellRec rb = rec_of(bit);
                    if (rb.id < a) continue;
                    double wv[4];
                    contact_rec(P, ax, ay, az, a, rb, rmp, wv);
                    fx = fx + wv[0];
                    fy = fy + wv[1];
                    fz = fz + wv[2];
                    if (tau_on) tau += wv[3];
                }
            }
            const double fo[3] = {fx, fy, fz};
            finish(a, ax, ay, az, fo, tau, q);
        }
    }
    // argmax-y partial of this workgroup's particles for the next stage
    const MarkC r = block_mark(mine, smk);
    if (tid == 0) nxt.mp[wg] = MarkPart{r.x, r.y, r.z, r.i, 0};
}

// workgroups of k_stage (and of k_cells_append) for a slab
inline dim3 contact_tile_grid(const PartParams &P)
{
    return dim3((unsigned)tiles_of(P.ncx, TILE_X), (unsigned)tiles_of(P.ncy, TILE_Y),
                (unsigned)(tiles_of(P.ncz, TILE_Z) * P.nbox));
}

// ---------------------------------------------------------------------------
// Spread nodes (cluster.cpp:344-410) per grid line
// ---------------------------------------------------------------------------
constexpr int NODE_LANES = 16;                 // lanes per grid line
constexpr int NODE_LPB = 256 / NODE_LANES;     // lines per workgroup
constexpr int NODE_MAXE = 2 * LINE_CAP;        // entries per line on the fast path
constexpr int NODE_REGION = 32 * NODE_LPB;     // pool nodes per workgroup (32 per line)
constexpr int NODE_PER = NODE_MAXE / NODE_LANES;  // entries per lane

// Nodes of the spread pool: every workgroup of the node kernel owns a
// region of NODE_REGION nodes; lines that do not fit their workgroup's
// region, and overflowing lines, take nodes from the shared tail.
__host__ __device__ constexpr size_t spread_pool_nodes(int nlines, int n)
{
    return (size_t)((nlines + NODE_LPB - 1) / NODE_LPB) * NODE_REGION + 8 * (size_t)(n > 0 ? n : 1);
}

// Per line: the entries of its particles sorted by (x, id), summed per node
// in ascending id; node_base / node_cnt locate them in the pool and the
// line's counter is cleared.  Lines beyond LINE_CAP particles take an exact
// repeated-selection path over the overflow list.
// so (optional): the forces of the overflow path come from the particles'
// StepOut blocks (the step's stage 2) instead of f0, f1, f2.
template <typename T>
__global__ __launch_bounds__(256) void k_spread_nodes(PartParams P, LineTab lt, const double *__restrict__ px,
                                                      const double *__restrict__ py, const double *__restrict__ pz,
                                                      const double *__restrict__ f0, const double *__restrict__ f1,
                                                      const double *__restrict__ f2, const StepOut *__restrict__ so,
                                                      SpreadNode<T> *__restrict__ nodes, int nlines)
{
    constexpr int W = NODE_LANES;
    __shared__ uint64_t key[NODE_LPB][NODE_MAXE];
    __shared__ uint64_t skey[NODE_LPB][NODE_MAXE];
    __shared__ T sval[NODE_LPB][NODE_MAXE][3];
    __shared__ int hc[NODE_LPB];
    const int sub = threadIdx.x / W, lane = threadIdx.x % W;
    const int line = blockIdx.x * NODE_LPB + sub;
    const int shift = (int)(threadIdx.x & 63) - lane;
    const bool on = line < nlines;
    const int cnt = on ? lt.cnt[line] : 0;
    const bool fast = cnt <= LINE_CAP;
    const int lb = on ? line % P.nlines_box : 0;
    const int yl = lb % P.ny, zl = lb / P.ny;
    const int m = fast ? cnt : 0;
    const int ne = 2 * m;
    // this lane's entries (particles lane, lane + W): keys to LDS, values kept
    uint64_t mk[NODE_PER];
    T mv[NODE_PER][3];
#pragma unroll
    for (int t = 0; t < NODE_PER / 2; t++) {
        const int k = lane + t * W;
        if (k < m) {
            const LineEnt<T> e = reinterpret_cast<const LineEnt<T> *>(lt.ent)[(size_t)line * LINE_CAP + k];
            const int a = e.id, x0 = e.x0, x1 = (x0 + 1) % P.nx;
            const T *w0 = e.w0, *w1 = e.w1;
            mk[2 * t] = ((uint64_t)(uint32_t)x0 << 32) | (uint32_t)a;
            mk[2 * t + 1] = ((uint64_t)(uint32_t)x1 << 32) | (uint32_t)a;
            key[sub][2 * k] = mk[2 * t];
            key[sub][2 * k + 1] = mk[2 * t + 1];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                mv[2 * t][d] = w0[d];
                mv[2 * t + 1][d] = w1[d];
            }
        }
    }
    __syncthreads();
    // rank sort (keys are unique: a particle meets a line at two x)
#pragma unroll
    for (int t = 0; t < NODE_PER / 2; t++) {
        const int k = lane + t * W;
        if (k < m) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const uint64_t v = mk[2 * t + h];
                int r = 0;
                for (int j = 0; j < ne; j++) r += key[sub][j] < v ? 1 : 0;
                skey[sub][r] = v;
#pragma unroll
                for (int d = 0; d < 3; d++) sval[sub][r][d] = mv[2 * t + h][d];
            }
        }
    }
    __syncthreads();
    // nodes of each line (heads of runs of equal x), then the line's place
    // in the workgroup's region: a prefix sum over the workgroup's lines
    int heads = 0;
    for (int c0 = 0; c0 < ne; c0 += W) {
        const int i = c0 + lane;
        const bool head = i < ne && (i == 0 || (skey[sub][i - 1] >> 32) != (skey[sub][i] >> 32));
        heads += __popcll((__ballot(head) >> shift) & ((1ull << W) - 1ull));
    }
    if (lane == 0) hc[sub] = fast ? heads : 0;
    __syncthreads();
    if (!on) return;
    if (fast) {
        int off = 0;
        for (int k = 0; k < sub; k++) off += hc[k];
        int base;
        if (off + heads <= NODE_REGION) base = blockIdx.x * NODE_REGION + off;
        else {
            const int tail = ((nlines + NODE_LPB - 1) / NODE_LPB) * NODE_REGION;
            base = 0;
            if (lane == 0) base = tail + atomicAdd(lt.pool_top, heads);
            base = __shfl(base, 0, W);
        }
        int done = 0;
        for (int c0 = 0; c0 < ne; c0 += W) {
            const int i = c0 + lane;
            const bool head = i < ne && (i == 0 || (skey[sub][i - 1] >> 32) != (skey[sub][i] >> 32));
            const unsigned long long seg = (__ballot(head) >> shift) & ((1ull << W) - 1ull);
            if (head) {
                const uint32_t x = (uint32_t)(skey[sub][i] >> 32);
                T s0 = T(0), s1 = T(0), s2 = T(0);
                for (int j = i; j < ne && (uint32_t)(skey[sub][j] >> 32) == x; j++) {
                    s0 += sval[sub][j][0];
                    s1 += sval[sub][j][1];
                    s2 += sval[sub][j][2];
                }
                SpreadNode<T> nd;
                nd.x = (int)x;
                nd.v[0] = s0;
                nd.v[1] = s1;
                nd.v[2] = s2;
                nodes[base + done + __popcll(seg & ((1ull << lane) - 1ull))] = nd;
            }
            done += __popcll(seg);
        }
        if (lane == 0) {
            lt.node_base[line] = base;
            lt.node_cnt[line] = heads;
            lt.cnt[line] = 0;
        }
        return;
    }
    // overflow: the first lane walks the nodes in ascending x and, per
    // node, the particles in ascending id (table part and overflow list)
    if (lane != 0) return;
    const int tail = ((nlines + NODE_LPB - 1) / NODE_LPB) * NODE_REGION;
    auto visit = [&](auto &&f) {
        for (int k = 0; k < LINE_CAP; k++) f(reinterpret_cast<const LineEnt<T> *>(lt.ent)[(size_t)line * LINE_CAP + k].id);
        for (int k = lt.ovf_head[line]; k >= 0;) {
            const int2 e = lt.ovf[k];
            f(e.x);
            k = e.y;
        }
    };
    int nh = 0;
    for (int pass = 0; pass < 2; pass++) {
        int base = 0;
        if (pass == 1 && nh > 0) base = tail + atomicAdd(lt.pool_top, nh);
        int k = 0;
        int64_t last_x = -1;
        for (;;) {
            // next node: the smallest x above last_x
            int64_t nxn = INT64_MAX;
            visit([&](int a) {
                int x0, x1;
                T w0[3], w1[3];
                line_entries<T>(P, px[a], py[a], pz[a], so ? so[a].fpp[0] : f0[a], so ? so[a].fpp[1] : f1[a], so ? so[a].fpp[2] : f2[a], yl, zl, &x0, &x1, w0, w1);
                if (x0 > last_x && x0 < nxn) nxn = x0;
                if (x1 > last_x && x1 < nxn) nxn = x1;
            });
            if (nxn == INT64_MAX) break;
            if (pass == 1) {
                T sm[3] = {T(0), T(0), T(0)};
                int64_t last_id = -1;
                for (;;) {
                    int64_t nid = INT64_MAX;
                    T wn[3] = {T(0), T(0), T(0)};
                    visit([&](int a) {
                        if (a <= last_id || a >= nid) return;
                        int x0, x1;
                        T w0[3], w1[3];
                        line_entries<T>(P, px[a], py[a], pz[a], so ? so[a].fpp[0] : f0[a], so ? so[a].fpp[1] : f1[a], so ? so[a].fpp[2] : f2[a], yl, zl, &x0, &x1, w0, w1);
                        if (x0 == nxn) { nid = a; wn[0] = w0[0]; wn[1] = w0[1]; wn[2] = w0[2]; }
                        else if (x1 == nxn) { nid = a; wn[0] = w1[0]; wn[1] = w1[1]; wn[2] = w1[2]; }
                    });
                    if (nid == INT64_MAX) break;
                    sm[0] += wn[0];
                    sm[1] += wn[1];
                    sm[2] += wn[2];
                    last_id = nid;
                }
                SpreadNode<T> nd;
                nd.x = (int)nxn;
                nd.v[0] = sm[0];
                nd.v[1] = sm[1];
                nd.v[2] = sm[2];
                nodes[base + k] = nd;
            }
            k++;
            last_x = nxn;
        }
        if (pass == 0) nh = k;
        else {
            lt.node_base[line] = base;
            lt.node_cnt[line] = nh;
        }
    }
    lt.cnt[line] = 0;
    lt.ovf_head[line] = -1;
}

// The id-indexed particle arrays from the step's blocks: p_old and u_old
// from Carry, u, f_pp, f_pi, grad c, c and tau from StepOut (for the API,
// snapshots and diagnostics, after steps whose stages wrote only the blocks).
template <typename T>
__global__ __launch_bounds__(256) void k_unpack(PartState S, int n)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const Carry<T> &cy = reinterpret_cast<const Carry<T> *>(S.carry)[a];
    const StepOut &o = reinterpret_cast<const StepOut *>(S.out)[a];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        S.po[d][a] = cy.po[d];
        S.uo[d][a] = cy.uo[d];
        S.u[d][a] = o.u[d];
        S.fpp[d][a] = o.fpp[d];
        S.fpi[d][a] = o.fpi[d];
        S.dcdx[d][a] = o.dcdx[d];
    }
    S.cloc[a] = o.cloc;
    S.tau_i[a] = o.tau;
}

}  // namespace fipi
