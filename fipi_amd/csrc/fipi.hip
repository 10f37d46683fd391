// libfipi: host side of the MI355X FIPI time step behind the C ABI of
// include/fipi.h.  One context = one periodic box (or a slab of equal
// boxes) on the current HIP device; all state lives in device memory.
//
// Per step (reference src/main.cpp:75-222), with R = real grid, H = half
// spectrum scratch, K = persistent spectra:
//   (x fwd  conv = u.grad c          R(6)  -> H0, first step after set_state)
//   y fwd   conv                     H0    -> H0
//   z hub   CH SBDF2 update          H0,K  -> K(c,hist), H0,H1,H2  (main.cpp:95-140)
//           (c/G, dz c/G, lambda k^2 c/G)
//   y inv   c, dz c -> + dy c; lambda k^2 c                         (main.cpp:160-162)
//   x pass  c, grad c (R); f(c) fwd -> fk; mu = f + lambda(-lap c);
//           F = mu grad c + spread + buoyancy fwd -> H0..H2   (main.cpp:153-175)
//   y fwd   f, F
//   z hub   Stokes projection of F; f -> f_hat (K) for the next step (main.cpp:177-178)
//   y inv   u
//   x pass  u (R); conv = u.grad c fwd -> H0 for the next step  (main.cpp:179, 91-92)
//   two Heun particle stages + the next step's spread, on a second stream
//   beside the next step's first passes within one call  (main.cpp:198-217)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fipi.h"
#include "field_kernels.hpp"
#include "particle_kernels.hpp"
#include "particle_stage.hpp"

using namespace fipi;

namespace {

thread_local std::string g_last_error;

template <class F>
bool with_n(int n, F &&f)
{
    switch (n) {
    case 16: f(std::integral_constant<int, 16>{}); return true;
    case 32: f(std::integral_constant<int, 32>{}); return true;
    case 48: f(std::integral_constant<int, 48>{}); return true;
    case 64: f(std::integral_constant<int, 64>{}); return true;
    case 80: f(std::integral_constant<int, 80>{}); return true;
    case 96: f(std::integral_constant<int, 96>{}); return true;
    case 128: f(std::integral_constant<int, 128>{}); return true;
    case 256: f(std::integral_constant<int, 256>{}); return true;
    default: return false;
    }
}

// tile width in bytes (64 or 128) as a compile-time constant
template <class F>
void with_tb(int tb, F &&f)
{
    if (tb == 64) f(std::integral_constant<int, 64>{});
    else f(std::integral_constant<int, 128>{});
}

bool supported_n(int n)
{
    return n == 16 || n == 32 || n == 48 || n == 64 || n == 80 || n == 96 || n == 128 || n == 256;
}

}  // namespace

struct fipi_ctx {
    fipi_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool fp64 = true;
    bool has_state = false;
    int t = 0;
    int *t_dev = nullptr;  // device copy of t (read by kernels a captured graph replays)
    int nbox = 1;          // equal boxes in the slab; G and GH are per box
    // tile widths (bytes) of the y passes, the CH/mu/init hubs and the
    // Stokes hub; results do not depend on them (every line sees the same
    // arithmetic), only the speed does
    // and the values per thread (FFT plan) of the y passes, 0 = the default
    // of the size; that choice changes rounding, not the transform
    struct Tune {
        int y_tb = 128, hub_tb = 128, stokes_tb = 128, y_e = 0;
    } tune;
    Dims d{};
    size_t G = 0, GH = 0;
    size_t esz = 8;  // bytes per real element
    // field state
    void *c = nullptr, *g[3] = {}, *u[3] = {};
    void *h[4] = {};
    void *ck[2] = {}, *fk = nullptr, *hist = nullptr;
    int cur = 0;
    // h[0] already holds the next step's conv (X) and conv_part its FP64
    // block sums: the previous step's last pass formed u.grad c
    bool conv_ready = false;
    // Velocity kept in spectral form between steps (u_lazy): the last pass
    // of a step forms u only in registers (for u.grad c) and writes no real
    // u, which saves three real fields of HBM writes a step; it leaves
    // u_hat (after the y transforms) in h[3], h[1], h[2] and the next step's
    // conv in h[0].  u_stale: the real u arrays are behind; materialize_u
    // forms them before anything outside the step reads them (getters,
    // snapshots, the tuner).  Off where the particle stage interpolates u
    // and in the ensemble and adjoint contexts.
    bool u_lazy = false, u_stale = false;
    // the step's particle stages wrote the per-particle blocks (Carry,
    // StepOut) but not the id-indexed arrays the API reads; unpack_particles
    // forms them before anything outside the step reads them
    bool parts_stale = false;
    // calibration extension: wetting source W (real, added to f; null in the
    // reference physics) and particles held at their positions (no Heun
    // stages; their contact forces are spread every step unchanged)
    void *W = nullptr;
    bool fix_particles = false;
    // Within fipi_step_async the particle stage of one step, and the spread
    // of its forces for the next step, run on pstream beside the next
    // step's first passes, which touch no particle state.  The stage reads
    // c and grad c, so that next step writes them into the spare set
    // (c_alt, g_alt; single-box contexts with particles) and swaps.  The
    // main stream waits on ev_join before the force pass that reads the
    // spread (spread_ready: it is built).
    hipStream_t pstream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool p_pending = false, spread_ready = false;
    void *c_alt = nullptr, *g_alt[3] = {};
    double *kx = nullptr, *ky = nullptr, *kz = nullptr;
    void *twx = nullptr, *twy = nullptr, *twz = nullptr;
    // particles
    int n = 0;
    PartParams pp{};
    PartState ps{};
    int ncell = 0;
    int *cell_cnt = nullptr, *cell_start = nullptr, *cell_ids = nullptr, *cell_cursor = nullptr;
    int nlines = 0;
    int *line_cnt = nullptr, *line_start = nullptr, *line_cursor = nullptr, *node_cnt = nullptr;
    void *spr_tmp = nullptr;  // spread node pool (8n nodes)
    // the step's particle stages: three cell tables of fixed capacity in
    // rotation (TabSet) and the spread's line buckets
    TabSet tabs{};
    LineTab lines{};
    bool tables_ready = false;
    // scratch of the facade's complex transform (allocated on first use)
    void *fs_re = nullptr, *fs_im = nullptr, *fs_h1 = nullptr, *fs_h2 = nullptr;
    int *bsum = nullptr;       // scan block totals
    double *ymax = nullptr;    // argmax-y block partials
    int *yidx = nullptr;
    int *d_flags = nullptr;  // [0] error bits, [1] contact-order overflow, [3] scratch flag of set_state
    int *h_flags = nullptr;  // pinned
    double *d_red = nullptr; // reduction scratch
    // FP32-mixed: FP64 k = 0 mode [c(0) ring of 2, history(0)], conv(0)
    double *k0 = nullptr, *conv_part = nullptr, *conv0 = nullptr;
    std::vector<void *> allocs;
    // per-launch timing (fipi_profile_steps)
    struct ProfRec {
        const char *name;
        double bytes;
        hipEvent_t a, b;
    };
    bool prof_on = false;
    std::vector<ProfRec> prof;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
};

namespace {

int set_err(fipi_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    g_last_error = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                           \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return set_err(ctx, FIPI_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

int dalloc(fipi_ctx *ctx, void **ptr, size_t bytes)
{
    *ptr = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) return set_err(ctx, FIPI_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    ctx->allocs.push_back(*ptr);
    return FIPI_OK;
}

// Wave numbers, src/field.cpp:108-115: i*2.0*pi/L with i <= N/2 kept positive.
std::vector<double> wave_numbers(int n, double pi, double L)
{
    std::vector<double> k(n);
    for (int i = 0; i < n; i++) k[i] = (i <= n / 2 ? i : i - n) * 2.0 * pi / L;
    return k;
}

// Forward twiddles exp(-2 pi i j / N), j < N, long-double accurate.
template <typename T>
std::vector<T> twiddles(int n)
{
    std::vector<T> w(2 * n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int j = 0; j < n; j++) {
        const long double a = two_pi * (long double)j / (long double)n;
        w[2 * j] = (T)cosl(a);
        w[2 * j + 1] = (T)(-sinl(a));
    }
    return w;
}


// Timing scope around one launch site when profiling is on: HIP events on
// the context stream, plus the site's algorithmic bytes (each array the
// kernel must touch, read or written once).
hipEvent_t prof_event(fipi_ctx *ctx)
{
    if (ctx->ev_used == ctx->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->ev_pool.push_back(e);
    }
    return ctx->ev_pool[ctx->ev_used++];
}

struct ProfScope {
    fipi_ctx *ctx;
    size_t idx = (size_t)-1;
    ProfScope(fipi_ctx *c, const char *name, double bytes) : ctx(c)
    {
        if (!ctx->prof_on) return;
        fipi_ctx::ProfRec r{name, bytes, prof_event(ctx), prof_event(ctx)};
        if (!r.a || !r.b) return;
        (void)hipEventRecord(r.a, ctx->stream);
        idx = ctx->prof.size();
        ctx->prof.push_back(r);
    }
    ~ProfScope()
    {
        if (idx != (size_t)-1) (void)hipEventRecord(ctx->prof[idx].b, ctx->stream);
    }
};

// at least one block: every kernel bounds-checks its index
inline dim3 grid1(size_t n, int b) { return dim3((unsigned)(n == 0 ? 1 : (n + b - 1) / b)); }

// x-pass lines per workgroup: the largest divisor of ny*nz/2 (complex
// lines, 2 real lines each) that fits in the x-pass workgroup (XT_of)
template <typename T, int N, int EV = PlanE_for<T, N>::E>
int xlines(const Dims &d)
{
    const int cl = d.ny * d.nz / 2;
    int l = XShape<T, N, EV>::LPW;
    while (cl % l != 0) l--;
    return l;
}

// x-pass grid (one slab of workgroups per box) and workgroup
template <typename T, int N, int EV = PlanE_for<T, N>::E>
dim3 xgrid(const Dims &d, int nbox)
{
    return dim3((unsigned)(d.ny * d.nz / 2 / xlines<T, N, EV>(d)), 1, (unsigned)nbox);
}
template <typename T, int N, int EV = PlanE_for<T, N>::E>
dim3 xblock(const Dims &d)
{
    return dim3((unsigned)(xlines<T, N, EV>(d) * Line<T, N, EV>::P));
}

template <typename T>
int xblocks(const Dims &d)
{
    int r = 0;
    with_n(d.nx, [&](auto NC) { r = (int)xgrid<T, decltype(NC)::value>(d, 1).x; });
    return r;
}

// Exclusive scan of n counts into start[0..n] on the context stream;
// bsum holds ceil(n/1024) block totals.  Optionally reduces argmax-y
// partials into *mark.
void scan(fipi_ctx *ctx, const int *cnt, int *start, int n, int *bsum, const double *ymax, const int *yidx,
          int nparts, int *mark)
{
    const int nb = n > 0 ? (n + SCAN_ITEMS - 1) / SCAN_ITEMS : 1;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, ctx->stream, cnt, start, n, bsum);
    hipLaunchKernelGGL(k_scan_finish, dim3(nb), dim3(256), 0, ctx->stream, start, n, (const int *)bsum, nb, ymax, yidx,
                       nparts, mark);
}

// advance the device step counter (one thread)
__global__ void k_tick(int *t)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *t += 1;
}

// ---------------------------------------------------------------------------
// Templated launch sequences
// ---------------------------------------------------------------------------
template <typename T>
struct Impl {
    using C = cplx<T>;

    static T *R(void *p) { return static_cast<T *>(p); }
    static C *H(void *p) { return static_cast<C *>(p); }

    static KParams kparams(const fipi_ctx *ctx)
    {
        KParams kp;
        kp.dt = ctx->p.dt;
        kp.mobility = ctx->p.mobility;
        kp.lambda = ctx->p.lambda;
        kp.mu = ctx->p.mu;
        kp.inv_g = 1.0 / (double)((int64_t)ctx->p.nx * ctx->p.ny * ctx->p.nz);
        kp.inv_mu_g = ctx->p.mu != 0.0 ? kp.inv_g / ctx->p.mu : 0.0;  // u = 0 when mu = 0 (division by zero gives zero)
        kp.first = ctx->t == 0;
        kp.use_conv = ctx->p.switch_ns != 0;
        return kp;
    }

    static T fscale(const fipi_ctx *ctx) { return (T)(ctx->p.lambda / (ctx->p.epsilon * ctx->p.epsilon)); }

    template <int NF, class Prod>
    static void x_forward(fipi_ctx *ctx, Prod prod, C *o0, C *o1, double *part = nullptr)
    {
        const Dims d = ctx->d;
        with_n(d.nx, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            hipLaunchKernelGGL((k_x_forward<T, N, NF, Prod>), (xgrid<T, N>(d, ctx->nbox)), (xblock<T, N>(d)), 0, ctx->stream, prod, o0, o1, d,
                               (const C *)ctx->twx, part);
        });
    }

    template <int NF>
    static void x_inverse(fipi_ctx *ctx, const C *i0, const C *i1, const C *i2, T *o0, T *o1, T *o2)
    {
        const Dims d = ctx->d;
        with_n(d.nx, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            hipLaunchKernelGGL((k_x_inverse<T, N, NF>), (xgrid<T, N>(d, ctx->nbox)), (xblock<T, N>(d)), 0, ctx->stream, i0, i1, i2, o0, o1, o2, d,
                               (const C *)ctx->twx);
        });
    }

    // real u from the u_hat the last step left in h[3], h[1], h[2]; the same
    // inverse x transform (plan and arithmetic) as the step's last pass
    static void materialize_u(fipi_ctx *ctx)
    {
        if (!ctx->u_stale) return;
        const Dims d = ctx->d;
        with_n(d.nx, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            constexpr int XE = XE_conv<T, N>::v;
            hipLaunchKernelGGL((k_x_inverse<T, N, 3, XE>), (xgrid<T, N, XE>(d, ctx->nbox)), (xblock<T, N, XE>(d)), 0,
                               ctx->stream, (const C *)H(ctx->h[3]), (const C *)H(ctx->h[1]),
                               (const C *)H(ctx->h[2]), R(ctx->u[0]), R(ctx->u[1]), R(ctx->u[2]), d,
                               (const C *)ctx->twx);
        });
        ctx->u_stale = false;
    }

    // the id-indexed particle arrays from the last step's blocks, after the
    // particle stream has ended
    static void unpack_particles(fipi_ctx *ctx)
    {
        if (!ctx->parts_stale || ctx->n == 0) return;
        if (ctx->p_pending) (void)hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
        hipLaunchKernelGGL((k_unpack<T>), grid1(ctx->n, 256), dim3(256), 0, ctx->stream, ctx->ps, ctx->n);
        ctx->parts_stale = false;
    }

    template <int SIGN, int NF, bool GRAD>
    static void y_pass(fipi_ctx *ctx, C *a0, C *a1, C *a2, C *extra)
    {
        const Dims d = ctx->d;
        with_n(d.ny, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            with_tb(ctx->tune.y_tb, [&](auto TBC) {
                constexpr int TB = decltype(TBC)::value, TL = TileW<T, TB>::L;
                auto launch = [&](auto EC) {
                    constexpr int EY = decltype(EC)::value;
                    hipLaunchKernelGGL((k_y<T, N, SIGN, NF, GRAD, TB, EY>),
                                       dim3((unsigned)col_tiles(d.nxh, TL, TB), d.nz, ctx->nbox),
                                       dim3((TL * Line<T, N, EY>::P)), 0, ctx->stream, a0, a1, a2, extra, d,
                                       ctx->ky, (const C *)ctx->twy);
                };
                if constexpr (YE<T, N>::alt != 0) {
                    if (ctx->tune.y_e == YE<T, N>::alt) {
                        launch(std::integral_constant<int, YE<T, N>::alt>{});
                        return;
                    }
                }
                launch(std::integral_constant<int, YE<T, N>::v>{});
            });
        });
    }

    // the y-pass plans of this context's ny: {default, alternative or 0}
    static void y_plans(const fipi_ctx *ctx, int *def, int *alt)
    {
        *def = *alt = 0;
        with_n(ctx->d.ny, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            *def = YE<T, N>::v;
            *alt = YE<T, N>::alt;
        });
    }

    // z-pass (hub) grid and workgroup for tile width TB
    template <int TB>
    static dim3 zgrid(const fipi_ctx *ctx)
    {
        constexpr int TL = TileW<T, TB>::L;
        return dim3((unsigned)col_tiles(ctx->d.nxh, TL, TB), ctx->d.ny, ctx->nbox);
    }
    template <int TB, int N>
    static dim3 zblock()
    {
        return dim3(TileW<T, TB>::L * Line<T, N>::P);
    }

    template <int SIGN>
    static void z_pass(fipi_ctx *ctx, C *a)
    {
        const Dims d = ctx->d;
        with_n(d.nz, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            with_tb(ctx->tune.hub_tb, [&](auto TBC) {
                constexpr int TB = decltype(TBC)::value;
                hipLaunchKernelGGL((k_z<T, N, SIGN, TB>), zgrid<TB>(ctx), (zblock<TB, N>()), 0, ctx->stream, a, d,
                                   (const C *)ctx->twz);
            });
        });
    }

    // c^0 -> c_hat^0, f_hat^0 (K), grad c^0 (see set_state)
    static void init_fields(fipi_ctx *ctx)
    {
        const Dims d = ctx->d;
        C *h0 = H(ctx->h[0]), *h1 = H(ctx->h[1]), *h2 = H(ctx->h[2]);
        ProdCF<T> pr{R(ctx->c), fscale(ctx), (const T *)ctx->W};
        x_forward<2>(ctx, pr, h0, h1);
        y_pass<-1, 2, false>(ctx, h0, h1, nullptr, nullptr);
        const KParams kp = kparams(ctx);
        with_n(d.nz, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            with_tb(ctx->tune.hub_tb, [&](auto TBC) {
                constexpr int TB = decltype(TBC)::value;
                hipLaunchKernelGGL((k_hub_init<T, N, TB>), zgrid<TB>(ctx), (zblock<TB, N>()), 0, ctx->stream, h0, h1,
                                   H(ctx->ck[ctx->cur]), H(ctx->fk), kp, d, ctx->kz, (const C *)ctx->twz);
            });
        });
        y_pass<+1, 2, true>(ctx, h0, h1, nullptr, h2);
        cgrad_f(ctx, h2, h0);  // f (X) into h0 is not used
    }

    // c, grad c from c_hat (h0), dy c (dyc), dz c (h1); f (X) into fout
    static void cgrad_f(fipi_ctx *ctx, C *dyc, C *fout)
    {
        const Dims d = ctx->d;
        with_n(d.nx, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            hipLaunchKernelGGL((k_x_cgrad_f<T, N>), (xgrid<T, N>(d, ctx->nbox)), (xblock<T, N>(d)), 0, ctx->stream,
                               H(ctx->h[0]), dyc, H(ctx->h[1]), R(ctx->c), R(ctx->g[0]), R(ctx->g[1]), R(ctx->g[2]),
                               fout, fscale(ctx), d, ctx->kx, (const C *)ctx->twx, (const T *)ctx->W);
        });
    }

    // Spread of n forces (f0, f1, f2) at positions (px, py, pz): line
    // buckets, then the per-line node sums (node pool, line_start/node_cnt)
    static void spread_build(fipi_ctx *ctx, const PartParams &pp, const double *px, const double *py,
                             const double *pz, const double *f0, const double *f1, const double *f2,
                             const int *box)
    {
        hipStream_t st = ctx->stream;
        LineTab lt = ctx->lines;
        ctx->spread_ready = false;  // the pool now holds these nodes
        {
            ProfScope ps_(ctx, "spread_lines", 48.0 * pp.n + 4.0 * sizeof(LineEnt<T>) * pp.n);
            hipLaunchKernelGGL(k_tab_clear, dim3(1), dim3(64), 0, st, lt.cnt, 0, lt.ovf_cnt, (int *)nullptr,
                               lt.pool_top, (int *)nullptr);
            hipLaunchKernelGGL((k_spread_lines<T>), grid1(pp.n, 256), dim3(256), 0, st, pp, px, py, pz, f0, f1, f2,
                               box, lt, ctx->d_flags);
        }
        spread_nodes(ctx, pp, px, py, pz, f0, f1, f2, nullptr);
    }

    // so: forces of the overflow path from the StepOut blocks (after stage 2)
    static void spread_nodes(fipi_ctx *ctx, const PartParams &pp, const double *px, const double *py,
                             const double *pz, const double *f0, const double *f1, const double *f2,
                             const StepOut *so)
    {
        const int nl = ctx->nlines;
        // the four line entries of each particle in, its node sums out
        ProfScope ps_(ctx, "spread_nodes", 4.0 * sizeof(LineEnt<T>) * pp.n + 8.0 * pp.n * sizeof(SpreadNode<T>) + 12.0 * nl);
        hipLaunchKernelGGL((k_spread_nodes<T>), grid1(nl, NODE_LPB), dim3(256), 0, ctx->stream, pp, ctx->lines, px, py,
                           pz, f0, f1, f2, so, (SpreadNode<T> *)ctx->spr_tmp, nl);
    }

    // CSR cell list with ids ascending per cell (unit entries, adjoint)
    static void cells_build(fipi_ctx *ctx, const PartParams &pp, const PartState &ps, int *ids, int *start,
                            double *ymax, int *yidx)
    {
        hipStream_t st = ctx->stream;
        const double n = pp.n, nc = ctx->ncell;
        hipLaunchKernelGGL(k_cell_count, grid1(pp.n, 256), dim3(256), 0, st, pp, ps.p[0], ps.p[1], ps.p[2],
                           ps.box, ps.cell, ctx->cell_cnt, ps.err, ymax, yidx);
        {
            ProfScope ps_(ctx, "cells_scan_mark", 8.0 * nc + 12.0 * (pp.n / 256 + 1));
            scan(ctx, ctx->cell_cnt, start, ctx->ncell, ctx->bsum, ymax, yidx, (int)grid1(pp.n, 256).x, ps.mark);
        }
        {
            ProfScope ps_(ctx, "cells_scatter", 12.0 * n);
            hipLaunchKernelGGL(k_cell_scatter, grid1(pp.n, 256), dim3(256), 0, st, pp.n, ps.cell, start,
                               ctx->cell_cursor, ids);
        }
        {
            ProfScope ps_(ctx, "cells_sort", 4.0 * nc + 8.0 * n);
            hipLaunchKernelGGL(k_cell_sort, grid1(ctx->ncell, 256), dim3(256), 0, st, ctx->ncell, start, ids, ps,
                               ctx->cell_cnt, ctx->cell_cursor);
        }
    }

    // the main stream waits for the particle stage on pstream
    static void join_particles(fipi_ctx *ctx)
    {
        if (!ctx->p_pending) return;
        (void)hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
        ctx->p_pending = false;
    }

    // spread of the particle forces at the current positions (the next
    // force pass adds them)
    static void spread_now(fipi_ctx *ctx)
    {
        const PartState &ps = ctx->ps;
        spread_build(ctx, ctx->pp, ps.p[0], ps.p[1], ps.p[2], ps.fpp[0], ps.fpp[1], ps.fpp[2], ps.box);
    }

    // the particle stage on pstream from here (one stage pending at a time)
    static void fork_stage(fipi_ctx *ctx)
    {
        join_particles(ctx);
        particle_stages_beside(ctx);
    }

    // both Heun stages, the step counter and the next step's spread on
    // pstream, after the work queued so far on the main stream
    static void particle_stages_beside(fipi_ctx *ctx)
    {
        hipStream_t main = ctx->stream;
        (void)hipEventRecord(ctx->ev_fork, main);
        (void)hipStreamWaitEvent(ctx->pstream, ctx->ev_fork, 0);
        ctx->stream = ctx->pstream;  // the stage's launches go to pstream
        particle_stage(ctx, 1);
        particle_stage(ctx, 2);
        hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, ctx->pstream, ctx->t_dev);
        (void)hipEventRecord(ctx->ev_join, ctx->pstream);
        ctx->stream = main;
        ctx->p_pending = true;
    }

    static double ramp(double t, double thr)
    {
        // primitive.cpp:122-129
        if (t < 0) return 0.0;
        if (t < thr) return t / thr;
        return 1.0;
    }

    // The table the next stage reads (stage 1 of step t: (2t) % 3) from the
    // current positions, after the state was set or restored, with one
    // argmax partial per stage workgroup; the table that stage appends to is
    // emptied too (no stage before it did).
    static void tables_init(fipi_ctx *ctx)
    {
        const PartParams &pp = ctx->pp;
        const PartState &ps = ctx->ps;
        const CellTab &T0 = ctx->tabs.t[(2 * ctx->t) % 3], &T1 = ctx->tabs.t[(2 * ctx->t + 1) % 3];
        for (const CellTab *tb : {&T0, &T1})
            hipLaunchKernelGGL(k_tab_clear, grid1(ctx->ncell, 256), dim3(256), 0, ctx->stream, tb->cnt, ctx->ncell,
                               tb->ovf_cnt, (int *)nullptr, (int *)nullptr, tb->ovf_head);
        const dim3 tg = contact_tile_grid(pp);
        hipLaunchKernelGGL(k_cells_append, dim3(tg.x * tg.y * tg.z), dim3(256), 0, ctx->stream, pp,
                           (const double *)ps.p[0], (const double *)ps.p[1], (const double *)ps.p[2], ps.box, T0,
                           ctx->d_flags);
        ctx->tables_ready = true;
    }

    // One Heun stage (main.cpp:198-217) in one launch: interpolation, f_pi,
    // contacts, advection and the next sort, tile by tile over the stage's
    // cell table (TabSet); stage 2 also appends the spread lines of f_pp at
    // the new positions, whose node sums follow for the next step's force
    // pass (when the Stokes solve is on).
    static void particle_stage(fipi_ctx *ctx, int stage)
    {
        hipStream_t st = ctx->stream;
        const PartParams &pp = ctx->pp;
        const PartState &ps = ctx->ps;
        if (!ctx->tables_ready) tables_init(ctx);
        const bool ns = ctx->p.switch_ns != 0;
        const bool spread = stage == 2 && ns;
        const double n = pp.n;
        const StageFields<T> F{(const T *)R(ctx->c), (const T *)R(ctx->g[0]), (const T *)R(ctx->g[1]),
                               (const T *)R(ctx->g[2]), (const T *)R(ctx->u[0]), (const T *)R(ctx->u[1]),
                               (const T *)R(ctx->u[2])};
        const dim3 grid = contact_tile_grid(pp);
        if (pp.interp_u || pp.brown_amp > 0.0)
            hipLaunchKernelGGL((k_fluid_noise<T>), grid1(pp.n, 256), dim3(256), 0, st, pp, ps, F,
                               (const int *)ctx->t_dev, stage);
        {
            // per particle: its record, 32 gathered grid values, radius and
            // contact angle, the appended record; stage 1 writes p_old, u_old
            // and the new position, stage 2 reads p_old, u_old and writes the
            // step's state (p, u, f_pp, tau, c, grad c, f_pi) and 4 line slots
            const double per = 32.0 + 32.0 * sizeof(T) + 16.0 + 32.0 +
                               (stage == 1 ? 72.0 : 48.0 + 17 * 8.0 + (spread ? 16.0 : 0.0));
            ProfScope ps_(ctx, stage == 1 ? "stage_1" : "stage_2", n * per);
            if (stage == 1)
                hipLaunchKernelGGL((k_stage<1, T>), grid, dim3(256), 0, st, pp, ps, F, ctx->tabs, ctx->ncell,
                                   ctx->lines, ns ? 1 : 0, (const int *)ctx->t_dev, ctx->p.ramp_t, ctx->d_flags + 1);
            else
                hipLaunchKernelGGL((k_stage<2, T>), grid, dim3(256), 0, st, pp, ps, F, ctx->tabs, ctx->ncell,
                                   ctx->lines, spread ? 1 : 0, (const int *)ctx->t_dev, ctx->p.ramp_t,
                                   ctx->d_flags + 1);
        }
        if (spread) {
            spread_nodes(ctx, pp, ps.p[0], ps.p[1], ps.p[2], ps.fpp[0], ps.fpp[1], ps.fpp[2],
                         (const StepOut *)ps.out);
            ctx->spread_ready = true;
        }
        if (stage == 2) ctx->parts_stale = true;
    }

    // beside: this step's particle stage runs on pstream after the step's
    // passes, beside the next step's first passes, until the next step's
    // force pass (which reads its spread); join_end: the main stream waits
    // for it at the end of this step (the last step of a call)
    static void step(fipi_ctx *ctx, bool beside = false, bool join_end = true)
    {
        const Dims d = ctx->d;
        hipStream_t st = ctx->stream;
        const bool ns = ctx->p.switch_ns != 0;
        const bool parts = ctx->n > 0;
        C *h0 = H(ctx->h[0]), *h1 = H(ctx->h[1]), *h2 = H(ctx->h[2]), *h3 = H(ctx->h[3]), *fk = H(ctx->fk);
        // algorithmic bytes of one real field and one half spectrum (nx/2+1
        // columns; the padding of the row pitch is not counted)
        const double Rb = (double)ctx->G * ctx->nbox * sizeof(T);
        const double Hb = (double)d.nxh * d.ny * d.nz * ctx->nbox * 2 * sizeof(T);
        const bool mixed = !std::is_same<T, double>::value;
        const bool fork = beside && parts && !ctx->fix_particles;
        if (ns) {
            if (!ctx->conv_ready) {
                materialize_u(ctx);
                ProdConv<T> pr{R(ctx->u[0]), R(ctx->u[1]), R(ctx->u[2]), R(ctx->g[0]), R(ctx->g[1]), R(ctx->g[2])};
                ProfScope ps_(ctx, "x_fwd_conv", 6 * Rb + Hb);
                x_forward<1>(ctx, pr, h0, nullptr, mixed ? ctx->conv_part : nullptr);
            }
            ctx->conv_ready = false;
            ProfScope ps_(ctx, "y_fwd_1", 2 * Hb);
            y_pass<-1, 1, false>(ctx, h0, nullptr, nullptr, nullptr);
        }
        const KParams kp = kparams(ctx);
        const int nxt = ctx->cur ^ 1;
        {
            // reads conv, f_hat^n, c^n, history; writes c^{n+1}, history,
            // c/G, dz c/G and (with Stokes) lambda k^2 c/G
            ProfScope ps_(ctx, "z_hub_ch", (ns ? 9 : 7) * Hb);
            with_n(d.nz, [&](auto NC) {
                constexpr int N = decltype(NC)::value;
                with_tb(ctx->tune.hub_tb, [&](auto TBC) {
                    constexpr int TB = decltype(TBC)::value;
                    // FP64 conv(0) = sum of conv over the grid, fixed order, in the hub
                    const K0 k0{mixed ? ctx->k0 : nullptr, ctx->conv0, ctx->conv_part, xblocks<T>(d), ctx->cur};
                    if (kp.first)
                        hipLaunchKernelGGL((k_hub_ch<T, N, TB, true>), zgrid<TB>(ctx), (zblock<TB, N>()), 0, st, h0,
                                           h1, ns ? h2 : nullptr, (const C *)ctx->ck[ctx->cur], H(ctx->ck[nxt]),
                                           (const C *)fk, H(ctx->hist), kp, d, ctx->kx, ctx->ky, ctx->kz,
                                           (const C *)ctx->twz, k0);
                    else
                        hipLaunchKernelGGL((k_hub_ch<T, N, TB, false>), zgrid<TB>(ctx), (zblock<TB, N>()), 0, st, h0,
                                           h1, ns ? h2 : nullptr, (const C *)ctx->ck[ctx->cur], H(ctx->ck[nxt]),
                                           (const C *)fk, H(ctx->hist), kp, d, ctx->kx, ctx->ky, ctx->kz,
                                           (const C *)ctx->twz, k0);
                });
            });
        }
        ctx->cur = nxt;
        {
            ProfScope ps_(ctx, "y_inv_grad", 5 * Hb);
            y_pass<+1, 2, true>(ctx, h0, h1, nullptr, h3);
        }
        if (ns) {
            ProfScope ps_(ctx, "y_inv_1", 2 * Hb);
            y_pass<+1, 1, false>(ctx, h2, nullptr, nullptr, nullptr);
        }
        if (ctx->p_pending) {
            // the previous step's particle stage still reads c and grad c
            std::swap(ctx->c, ctx->c_alt);
            for (int a = 0; a < 3; a++) std::swap(ctx->g[a], ctx->g_alt[a]);
        }
        if (!ns) {
            {
                ProfScope ps_(ctx, "x_inv_cgrad_f", 4 * Hb + 4 * Rb);
                cgrad_f(ctx, h3, fk);
            }
            {
                ProfScope ps_(ctx, "y_fwd_f", 2 * Hb);
                y_pass<-1, 1, false>(ctx, fk, nullptr, nullptr, nullptr);
            }
            ProfScope ps_(ctx, "z_fwd_f", 2 * Hb);
            z_pass<-1>(ctx, fk);
        }
        if (ns) {
            join_particles(ctx);
            if (parts && !ctx->spread_ready) spread_now(ctx);
            ctx->spread_ready = false;
            {
                ProfScope ps_(ctx, "x_cgrad_force", 8 * Hb + 4 * Rb + (parts ? 8.0 * ctx->n * sizeof(SpreadNode<T>) : 0.0));
                with_n(d.nx, [&](auto NC) {
                    constexpr int N = decltype(NC)::value;
                    hipLaunchKernelGGL((k_x_cgrad_force<T, N>), (xgrid<T, N>(d, ctx->nbox)), (xblock<T, N>(d)), 0, st,
                                       h0, (const C *)h3, h1, h2, fk, R(ctx->c), R(ctx->g[0]), R(ctx->g[1]),
                                       R(ctx->g[2]), (const int *)ctx->line_start, (const int *)ctx->node_cnt,
                                       (const SpreadNode<T> *)ctx->spr_tmp, parts ? 1 : 0, fscale(ctx),
                                       (const T *)ctx->W, (T)ctx->p.rho1, (T)ctx->p.rho2, (T)ctx->p.gravity, d,
                                       ctx->kx, (const C *)ctx->twx);
                });
            }
            {
                ProfScope ps_(ctx, "y_fwd_f", 2 * Hb);
                y_pass<-1, 1, false>(ctx, fk, nullptr, nullptr, nullptr);
            }
            {
                ProfScope ps_(ctx, "y_fwd_3", 6 * Hb);
                y_pass<-1, 3, false>(ctx, h0, h1, h2, nullptr);
            }
            {
                // and f (XY) -> f_hat (K) for the next CH update
                ProfScope ps_(ctx, "z_hub_stokes", 8 * Hb);
                with_n(d.nz, [&](auto NC) {
                    constexpr int N = decltype(NC)::value;
                    with_tb(ctx->tune.stokes_tb, [&](auto TBC) {
                        constexpr int TB = decltype(TBC)::value;
                        hipLaunchKernelGGL((k_hub_stokes<T, N, TB>), zgrid<TB>(ctx), (zblock<TB, N>()), 0, st, h0,
                                           h1, h2, fk, kp, d, ctx->kx, ctx->ky, ctx->kz, (const C *)ctx->twz);
                    });
                });
            }
            {
                ProfScope ps_(ctx, "y_inv_3", 6 * Hb);
                y_pass<+1, 3, false>(ctx, h0, h1, h2, nullptr);
            }
            {
                // u^{n+1} and, for the next step, conv = u^{n+1}.grad c^{n+1}
                // (X) in h[0]; with u_lazy, conv goes to h[3] (free after the
                // force pass), no real u is written, and h[0] / h[3] swap
                // roles, so that u_hat stays in h[3], h[1], h[2]
                const bool lazy = ctx->u_lazy;
                ProfScope ps_(ctx, "x_inv_u_conv", 4 * Hb + (lazy ? 3 : 6) * Rb);
                with_n(d.nx, [&](auto NC) {
                    constexpr int N = decltype(NC)::value;
                    constexpr int XE = XE_conv<T, N>::v;
                    hipLaunchKernelGGL((k_x_inv_conv<T, N, XE>), (xgrid<T, N, XE>(d, ctx->nbox)), (xblock<T, N, XE>(d)), 0, st, h0, (const C *)h1,
                                       (const C *)h2, lazy ? h3 : h0, lazy ? nullptr : R(ctx->u[0]),
                                       lazy ? nullptr : R(ctx->u[1]), lazy ? nullptr : R(ctx->u[2]),
                                       (const T *)R(ctx->g[0]), (const T *)R(ctx->g[1]), (const T *)R(ctx->g[2]), d,
                                       (const C *)ctx->twx, mixed ? ctx->conv_part : nullptr);
                });
                ctx->conv_ready = true;
                if (lazy) {
                    std::swap(ctx->h[0], ctx->h[3]);
                    ctx->u_stale = true;
                }
            }
        }
        if (fork) {
            // beside the next step's first passes, which touch no particle
            // state (measured: starting it right after the force pass, beside
            // the Stokes passes, is slower)
            fork_stage(ctx);
            if (join_end) join_particles(ctx);
        } else {
            join_particles(ctx);
            if (parts && !ctx->fix_particles) {
                particle_stage(ctx, 1);
                particle_stage(ctx, 2);
            }
            hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, st, ctx->t_dev);
        }
        ctx->t++;
    }

    // unit entries: the caller's half spectrum is unpadded ([nz][ny][nx/2+1])
    struct Unpadded {
        fipi_ctx *ctx;
        int saved;
        explicit Unpadded(fipi_ctx *c) : ctx(c), saved(c->d.nxp) { ctx->d.nxp = ctx->d.nxh; }
        ~Unpadded() { ctx->d.nxp = saved; }
    };

    // unit entry: R2C of one field
    static void r2c(fipi_ctx *ctx, const T *in, C *out)
    {
        Unpadded u(ctx);
        x_forward<1>(ctx, ProdCopy<T>{in}, out, nullptr);
        y_pass<-1, 1, false>(ctx, out, nullptr, nullptr, nullptr);
        z_pass<-1>(ctx, out);
    }

    static void c2r(fipi_ctx *ctx, C *in, T *out)
    {
        Unpadded u(ctx);
        z_pass<+1>(ctx, in);
        y_pass<+1, 1, false>(ctx, in, nullptr, nullptr, nullptr);
        x_inverse<1>(ctx, in, nullptr, nullptr, out, nullptr, nullptr);
    }
};

// scale a real array (c2r normalisation in the unit entry)
template <typename T>
__global__ void k_scale(T *a, size_t n, double s)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (T)((double)a[i] * s);
}

template <typename T>
__global__ void k_to_double(const T *a, double *b, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

template <typename T>
__global__ void k_from_double(const double *a, T *b, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (T)a[i];
}

__global__ void k_aos_to_soa(const double *a, double *x, double *y, double *z, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { x[i] = a[3 * i]; y[i] = a[3 * i + 1]; z[i] = a[3 * i + 2]; }
}

__global__ void k_soa_to_aos(const double *x, const double *y, const double *z, double *a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { a[3 * i] = x[i]; a[3 * i + 1] = y[i]; a[3 * i + 2] = z[i]; }
}

template <typename T>
__global__ void k_interp_field(PartParams P, const T *grid, const double *x, const double *y, const double *z,
                               double *out)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.n) return;
    out[a] = grid_cell_ok(P, x[a], y[a], z[a]) ? interp1(P, grid, x[a], y[a], z[a]) : 0.0;
}

template <typename T>
__global__ void k_interp2_field(PartParams P, const T *grid, const double *x, const double *y, const double *z,
                                double *out)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.n) return;
    out[a] = grid_cell_ok(P, x[a], y[a], z[a]) ? interp2(P, grid, x[a], y[a], z[a]) : 0.0;
}

// element::get_f_rand of n particles (AoS out)
__global__ void k_brownian_unit(PartParams P, int t, int stage, double *out)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.n) return;
    for (int d = 0; d < 3; d++) out[3 * a + d] = P.brown_amp * brownian_g(P.noise_seed, t, stage, a, d, P.pi);
}

// element::get_f_drag (element.cpp:147-151): -6 pi r mu (u - uf), AoS out
__global__ void k_drag_unit(PartParams P, PartState S, double *out, int n)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const double c = -6 * P.pi * S.rad[a] * P.mu;
    for (int d = 0; d < 3; d++) out[3 * a + d] = c * (S.u[d][a] - S.uf[d][a]);
}

// Diagnostics partial sums: mean phi, free energy density, |grad phi|.
template <typename T>
__global__ __launch_bounds__(256) void k_diag_partial(const T *c, const T *g0, const T *g1, const T *g2, size_t G,
                                                      double lambda, double eps, double *part)
{
    __shared__ double s[3][256];
    double a = 0, b = 0, e = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < G; i += (size_t)gridDim.x * 256) {
        const double cv = c[i], x = g0[i], y = g1[i], z = g2[i];
        const double gg = (x * x + y * y) + z * z;
        const double w = cv * cv - 1.0;
        a += cv;
        b += lambda * (0.5 * gg + 0.25 / (eps * eps) * (w * w));
        e += sqrt(gg);
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    s[2][threadIdx.x] = e;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < 3; q++) s[q][threadIdx.x] += s[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; q++) part[3 * blockIdx.x + q] = s[q][0];
}

constexpr int DIAG_BLOCKS = 1024;

static_assert(sizeof(fipi_params) == 224, "fipi_params layout (the bindings mirror it)");

int check_params(const fipi_params *p)
{
    if (!p) return set_err(nullptr, FIPI_ERR_ARG, "params is NULL");
    if (!supported_n(p->nx) || !supported_n(p->ny) || !supported_n(p->nz))
        return set_err(nullptr, FIPI_ERR_UNSUPPORTED,
                       "grid sizes must each be one of 16, 32, 48, 64, 80, 96, 128, 256");
    if (p->switch_ch != 1) return set_err(nullptr, FIPI_ERR_UNSUPPORTED, "switch_ch must be 1");
    if (p->precision != FIPI_FP64 && p->precision != FIPI_FP32_MIXED)
        return set_err(nullptr, FIPI_ERR_ARG, "precision must be FIPI_FP64 or FIPI_FP32_MIXED");
    if (p->n_particles < 0) return set_err(nullptr, FIPI_ERR_ARG, "n_particles < 0");
    // the contact kernel orders partners by a 32-bit key of (cell, id)
    if (p->n_particles >= (1 << 26)) return set_err(nullptr, FIPI_ERR_UNSUPPORTED, "n_particles must be below 2^26");
    if (!(p->lx > 0 && p->ly > 0 && p->lz > 0 && p->dt > 0))
        return set_err(nullptr, FIPI_ERR_ARG, "box lengths and dt must be positive");
    if (p->interp_u < 0 || p->interp_u > 2) return set_err(nullptr, FIPI_ERR_ARG, "interp_u must be 0, 1 or 2");
    if (!(p->kt >= 0.0 && std::isfinite(p->kt))) return set_err(nullptr, FIPI_ERR_ARG, "kt must be finite and >= 0");
    if (p->switch_particle && p->n_particles > 0) {
        if (!(p->r_c > 0)) return set_err(nullptr, FIPI_ERR_ARG, "r_c must be positive");
        const double L[3] = {p->lx, p->ly, p->lz};
        for (int a = 0; a < 3; a++) {
            const int nc = (int)floor(L[a] / p->r_c + 2 * p->margin / p->r_c);
            if (nc < 3) return set_err(nullptr, FIPI_ERR_ARG, "fewer than 3 contact cells per axis");
        }
    }
    return FIPI_OK;
}

// Particle-kernel parameters of one box of p (a slab of nbox such boxes)
// with n particles in all; the contact grid is cluster.cpp:26-32.
PartParams part_params_of(const fipi_params &p, int n, int nbox)
{
    const double L[3] = {p.lx, p.ly, p.lz};
    int nc[3];
    double dc[3];
    for (int a = 0; a < 3; a++) {
        nc[a] = (int)floor(L[a] / p.r_c + 2 * p.margin / p.r_c);
        dc[a] = fmod(L[a] + 2 * p.margin, p.r_c) / nc[a] + p.r_c;
    }
    PartParams pp{};
    pp.n = n;
    pp.nx = p.nx; pp.ny = p.ny; pp.nz = p.nz;
    pp.dx = p.lx / p.nx; pp.dy = p.ly / p.ny; pp.dz = p.lz / p.nz;
    pp.lx = p.lx; pp.ly = p.ly; pp.lz = p.lz;
    pp.ncx = nc[0]; pp.ncy = nc[1]; pp.ncz = nc[2];
    pp.dcx = dc[0]; pp.dcy = dc[1]; pp.dcz = dc[2];
    pp.margin = p.margin;
    pp.r_p = p.r_p; pp.r_eq = p.r_eq; pp.k_n = p.k_n;
    pp.pi = p.pi; pp.mu = p.mu; pp.dt = p.dt; pp.epsilon = p.epsilon; pp.sigma = p.sigma; pp.a_pi = p.a_pi;
    pp.tanh2 = p.tanh2;
    pp.sqrt2 = std::pow(2.0, 0.5);
    pp.has_ctheta = 0;
    // uf needs the Stokes solve (cluster.cpp:313); element.cpp:139-143
    pp.interp_u = p.switch_ns ? p.interp_u : 0;
    pp.brown_amp = p.kt > 0.0 ? std::sqrt(12.0 * p.pi * p.r_p * p.mu * p.kt / p.dt) : 0.0;
    pp.noise_seed = p.noise_seed;
    pp.nbox = nbox;
    pp.ncell_box = (nc[0] > 0 && nc[1] > 0 && nc[2] > 0) ? nc[0] * nc[1] * nc[2] : 1;
    pp.nlines_box = p.ny * p.nz;
    pp.rs = (size_t)p.nx * p.ny * p.nz;
    return pp;
}

template <typename T>
int setup_buffers(fipi_ctx *ctx)
{
    const fipi_params &p = ctx->p;
    ctx->esz = sizeof(T);
    // FP32: half-width Stokes tiles (three fields in flight) double the
    // workgroups per CU
    if (sizeof(T) == 4) ctx->tune.stokes_tb = 64;
    const size_t nb = (size_t)ctx->nbox;
    const size_t G = ctx->G * nb, GH = ctx->GH * nb;  // whole slab
    int rc;
#define ALLOC(ptr, bytes) \
    if ((rc = dalloc(ctx, (void **)&(ptr), (bytes))) != FIPI_OK) return rc
    ALLOC(ctx->c, G * sizeof(T));
    if (ctx->n > 0 && nb == 1) {
        ALLOC(ctx->c_alt, G * sizeof(T));
        for (int a = 0; a < 3; a++) ALLOC(ctx->g_alt[a], G * sizeof(T));
    }
    for (int a = 0; a < 3; a++) {
        ALLOC(ctx->g[a], G * sizeof(T));
        ALLOC(ctx->u[a], G * sizeof(T));
        ALLOC(ctx->h[a], GH * 2 * sizeof(T));
    }
    ALLOC(ctx->h[3], GH * 2 * sizeof(T));
    ALLOC(ctx->ck[0], GH * 2 * sizeof(T));
    ALLOC(ctx->ck[1], GH * 2 * sizeof(T));
    ALLOC(ctx->fk, GH * 2 * sizeof(T));
    ALLOC(ctx->hist, GH * 2 * sizeof(T));
    ALLOC(ctx->kx, sizeof(double) * p.nx);
    ALLOC(ctx->ky, sizeof(double) * p.ny);
    ALLOC(ctx->kz, sizeof(double) * p.nz);
    ALLOC(ctx->twx, sizeof(T) * 2 * p.nx);
    ALLOC(ctx->twy, sizeof(T) * 2 * p.ny);
    ALLOC(ctx->twz, sizeof(T) * 2 * p.nz);
    ALLOC(ctx->d_flags, sizeof(int) * 4);
    ALLOC(ctx->t_dev, sizeof(int));
    ALLOC(ctx->d_red, sizeof(double) * (3 * DIAG_BLOCKS + 16));
    ALLOC(ctx->k0, sizeof(double) * K0_STRIDE * nb);
    ALLOC(ctx->conv0, sizeof(double) * (nb + 1));
    ALLOC(ctx->conv_part, sizeof(double) * nb * (size_t)(xblocks<T>(ctx->d) + 1));
    auto kx = wave_numbers(p.nx, p.pi, p.lx), ky = wave_numbers(p.ny, p.pi, p.ly), kz = wave_numbers(p.nz, p.pi, p.lz);
    HIPCHK(ctx, hipMemcpy(ctx->kx, kx.data(), sizeof(double) * p.nx, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->ky, ky.data(), sizeof(double) * p.ny, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->kz, kz.data(), sizeof(double) * p.nz, hipMemcpyHostToDevice));
    auto tx = twiddles<T>(p.nx), ty = twiddles<T>(p.ny), tz = twiddles<T>(p.nz);
    HIPCHK(ctx, hipMemcpy(ctx->twx, tx.data(), sizeof(T) * 2 * p.nx, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->twy, ty.data(), sizeof(T) * 2 * p.ny, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(ctx->twz, tz.data(), sizeof(T) * 2 * p.nz, hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemset(ctx->d_flags, 0, sizeof(int) * 4));
    HIPCHK(ctx, hipHostMalloc((void **)&ctx->h_flags, sizeof(int) * 4, hipHostMallocDefault));
    PartParams &pp = ctx->pp;
    pp = part_params_of(p, ctx->n, ctx->nbox);
    ctx->ncell = pp.ncell_box * ctx->nbox;
    ctx->nlines = pp.nlines_box * ctx->nbox;
    ALLOC(ctx->cell_cnt, sizeof(int) * (ctx->ncell + 1));
    ALLOC(ctx->cell_start, sizeof(int) * (ctx->ncell + 1));
    ALLOC(ctx->cell_cursor, sizeof(int) * (ctx->ncell + 1));
    ALLOC(ctx->line_cnt, sizeof(int) * (ctx->nlines + 1));
    ALLOC(ctx->line_start, sizeof(int) * (ctx->nlines + 1));
    ALLOC(ctx->line_cursor, sizeof(int) * (ctx->nlines + 1));
    ALLOC(ctx->node_cnt, sizeof(int) * (ctx->nlines + 1));
    HIPCHK(ctx, hipMemset(ctx->line_start, 0, sizeof(int) * (ctx->nlines + 1)));
    // counters start at zero; each build's last kernel clears what it used
    HIPCHK(ctx, hipMemset(ctx->cell_cnt, 0, sizeof(int) * (ctx->ncell + 1)));
    HIPCHK(ctx, hipMemset(ctx->cell_cursor, 0, sizeof(int) * (ctx->ncell + 1)));
    HIPCHK(ctx, hipMemset(ctx->line_cnt, 0, sizeof(int) * (ctx->nlines + 1)));
    HIPCHK(ctx, hipMemset(ctx->line_cursor, 0, sizeof(int) * (ctx->nlines + 1)));
    const size_t n = ctx->n > 0 ? ctx->n : 1;
    ALLOC(ctx->cell_ids, sizeof(int) * n);
    ALLOC(ctx->spr_tmp, sizeof(SpreadNode<T>) * spread_pool_nodes(ctx->nlines, (int)n));
    // the step's cell tables and spread line buckets
    for (int k = 0; k < 3; k++) {
        CellTab &tb = ctx->tabs.t[k];
        ALLOC(tb.cnt, sizeof(int) * (ctx->ncell + 1));
        ALLOC(tb.rec, sizeof(CellRec) * CELL_CAP * (size_t)ctx->ncell);
        ALLOC(tb.ovf_cnt, sizeof(int));
        ALLOC(tb.ovf, sizeof(CellRec) * n);
        ALLOC(tb.ovf_head, sizeof(int) * (ctx->ncell + 1));
        HIPCHK(ctx, hipMemset(tb.ovf_head, 0xff, sizeof(int) * (ctx->ncell + 1)));  // -1: no chain
        {
            // argmax partials: one per stage workgroup
            const dim3 tg = contact_tile_grid(pp);
            ALLOC(tb.mp, sizeof(MarkPart) * ((size_t)tg.x * tg.y * tg.z));
        }
        HIPCHK(ctx, hipMemset(tb.cnt, 0, sizeof(int) * (ctx->ncell + 1)));
        HIPCHK(ctx, hipMemset(tb.ovf_cnt, 0, sizeof(int)));
    }
    {
        LineTab &lt = ctx->lines;
        lt.cnt = ctx->line_cnt;
        ALLOC(lt.ent, sizeof(LineEnt<T>) * LINE_CAP * (size_t)ctx->nlines);
        ALLOC(lt.ovf_cnt, sizeof(int));
        ALLOC(lt.ovf, sizeof(int2) * 4 * n);
        ALLOC(lt.ovf_head, sizeof(int) * (ctx->nlines + 1));
        HIPCHK(ctx, hipMemset(lt.ovf_head, 0xff, sizeof(int) * (ctx->nlines + 1)));  // -1: no chain
        lt.node_base = ctx->line_start;
        lt.node_cnt = ctx->node_cnt;
        ALLOC(lt.pool_top, sizeof(int));
        HIPCHK(ctx, hipMemset(lt.ovf_cnt, 0, sizeof(int)));
        HIPCHK(ctx, hipMemset(lt.pool_top, 0, sizeof(int)));
        HIPCHK(ctx, hipMemset(ctx->node_cnt, 0, sizeof(int) * (ctx->nlines + 1)));
    }
    {
        const size_t m = std::max<size_t>(std::max<size_t>(ctx->ncell, ctx->nlines), n);
        ALLOC(ctx->bsum, sizeof(int) * (m / SCAN_ITEMS + 2));
        ALLOC(ctx->ymax, sizeof(double) * (n / 256 + 2));
        ALLOC(ctx->yidx, sizeof(int) * (n / 256 + 2));
    }
    PartState &ps = ctx->ps;
    for (int a = 0; a < 3; a++) {
        ALLOC(ps.p[a], sizeof(double) * n);
        ALLOC(ps.po[a], sizeof(double) * n);
        ALLOC(ps.u[a], sizeof(double) * n);
        ALLOC(ps.uo[a], sizeof(double) * n);
        ALLOC(ps.fpp[a], sizeof(double) * n);
        ALLOC(ps.fpi[a], sizeof(double) * n);
        ALLOC(ps.dcdx[a], sizeof(double) * n);
        ALLOC(ps.sp[a], sizeof(double) * n);
        ALLOC(ps.uf[a], sizeof(double) * n);
        ALLOC(ps.frand[a], sizeof(double) * n);
    }
    ALLOC(ps.rad, sizeof(double) * n);
    ALLOC(ps.cth, sizeof(double) * n);
    ALLOC(ps.cloc, sizeof(double) * n);
    ALLOC(ps.tau_i, sizeof(double) * n);
    ALLOC(ps.carry, sizeof(Carry<T>) * n);
    ALLOC(ps.out, sizeof(StepOut) * n);
    ALLOC(ps.cell, sizeof(int) * n);
    if (ctx->nbox > 1) {
        int *bx;
        ALLOC(bx, sizeof(int) * n);
        ps.box = bx;
    }
    ps.err = ctx->d_flags;
    ALLOC(ps.mark, sizeof(int));
    ALLOC(ps.mark_on, sizeof(int));
    HIPCHK(ctx, hipMemset(ps.mark_on, 0, sizeof(int)));
#undef ALLOC
    return FIPI_OK;
}

int copy_in_real(fipi_ctx *ctx, void *dst, const double *src, size_t n)
{
    if (ctx->fp64) {
        HIPCHK(ctx, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        return FIPI_OK;
    }
    double *tmp;
    HIPCHK(ctx, hipMalloc((void **)&tmp, n * sizeof(double)));
    HIPCHK(ctx, hipMemcpyAsync(tmp, src, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL((k_from_double<float>), grid1(n, 256), dim3(256), 0, ctx->stream, tmp, (float *)dst, n);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipFree(tmp));
    return FIPI_OK;
}

int copy_out_real(fipi_ctx *ctx, double *dst, const void *src, size_t n)
{
    if (ctx->fp64) {
        HIPCHK(ctx, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        return FIPI_OK;
    }
    double *tmp;
    HIPCHK(ctx, hipMalloc((void **)&tmp, n * sizeof(double)));
    hipLaunchKernelGGL((k_to_double<float>), grid1(n, 256), dim3(256), 0, ctx->stream, (const float *)src, tmp, n);
    HIPCHK(ctx, hipMemcpyAsync(dst, tmp, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipFree(tmp));
    return FIPI_OK;
}

int check_flags(fipi_ctx *ctx)
{
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_flags, ctx->d_flags, sizeof(int) * 2, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    if (ctx->h_flags[0] & 1) return set_err(ctx, FIPI_ERR_RANGE, "particle outside the contact-cell sorting range");
    if (ctx->h_flags[0] & 2) return set_err(ctx, FIPI_ERR_RANGE, "particle outside the grid");
    return FIPI_OK;
}

// Temporary SoA particle arrays for the unit entry points.
struct TmpParts {
    fipi_ctx *ctx;
    PartState ps{};
    std::vector<void *> mem;
    int n;
    int alloc(void **p, size_t b)
    {
        hipError_t e = hipMalloc(p, b ? b : 16);
        if (e != hipSuccess) return set_err(ctx, FIPI_ERR_HIP, "hipMalloc failed");
        mem.push_back(*p);
        return FIPI_OK;
    }
    int init(fipi_ctx *c, const double *d_pos, int nn)
    {
        ctx = c;
        n = nn;
        const size_t b = sizeof(double) * (n > 0 ? n : 1);
        int rc;
        for (int a = 0; a < 3; a++) {
            if ((rc = alloc((void **)&ps.p[a], b))) return rc;
            if ((rc = alloc((void **)&ps.fpp[a], b))) return rc;
            if ((rc = alloc((void **)&ps.fpi[a], b))) return rc;
            if ((rc = alloc((void **)&ps.sp[a], b))) return rc;
        }
        if ((rc = alloc((void **)&ps.tau_i, b))) return rc;
        if ((rc = alloc((void **)&ps.cell, sizeof(int) * (n > 0 ? n : 1)))) return rc;
        if ((rc = alloc((void **)&ps.mark, sizeof(int)))) return rc;
        ps.err = ctx->d_flags;
        HIPCHK(ctx, hipMemsetAsync(ps.mark, 0xff, sizeof(int), ctx->stream));
        for (int a = 0; a < 3; a++) HIPCHK(ctx, hipMemsetAsync(ps.fpi[a], 0, b, ctx->stream));
        if (n > 0)
            hipLaunchKernelGGL(k_aos_to_soa, grid1(n, 256), dim3(256), 0, ctx->stream, d_pos, ps.p[0], ps.p[1],
                               ps.p[2], n);
        return FIPI_OK;
    }
    ~TmpParts()
    {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        for (void *p : mem) (void)hipFree(p);
    }
};

// cells build on caller arrays (unit entries)
int unit_cells(fipi_ctx *ctx, TmpParts &tp, int *start, int *ids)
{
    PartParams pp = ctx->pp;
    pp.n = tp.n;
    double *ymax;
    int *yidx, rc;
    if ((rc = tp.alloc((void **)&ymax, sizeof(double) * (tp.n / 256 + 2)))) return rc;
    if ((rc = tp.alloc((void **)&yidx, sizeof(int) * (tp.n / 256 + 2)))) return rc;
    if (ctx->fp64) Impl<double>::cells_build(ctx, pp, tp.ps, ids, start, ymax, yidx);
    else Impl<float>::cells_build(ctx, pp, tp.ps, ids, start, ymax, yidx);
    return FIPI_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

}  // extern "C"

// the real velocity arrays up to date (they lag while u_stale)
static void ensure_real_u(fipi_ctx *ctx)
{
    if (ctx->fp64) Impl<double>::materialize_u(ctx);
    else Impl<float>::materialize_u(ctx);
}

// the id-indexed particle arrays up to date (they lag while parts_stale)
static void ensure_particles(fipi_ctx *ctx)
{
    if (ctx->fp64) Impl<double>::unpack_particles(ctx);
    else Impl<float>::unpack_particles(ctx);
}

// A context of nbox equal boxes (the slab) with n_total particles in all.
static int create_ctx(const fipi_params *params, int nbox, int n_total, fipi_ctx **out)
{
    if (!out) return set_err(nullptr, FIPI_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_params(params);
    if (rc != FIPI_OK) return rc;
    if (nbox < 1) return set_err(nullptr, FIPI_ERR_ARG, "nbox < 1");
    if (nbox > 1 && (params->interp_u != 0 || params->kt != 0.0))
        return set_err(nullptr, FIPI_ERR_UNSUPPORTED, "interp_u and kt need a single-box context");
    if (n_total < 0 || n_total >= (1 << 26))
        return set_err(nullptr, FIPI_ERR_UNSUPPORTED, "the particle count must be below 2^26");
    fipi_ctx *ctx = new fipi_ctx();
    ctx->p = *params;
    ctx->nbox = nbox;
    ctx->fp64 = params->precision == FIPI_FP64;
    {
        const int nxh = params->nx / 2 + 1;
        // row pitch: whole 128-byte lines, plus one spare line (measured at
        // 256^3: the z hubs 4-6 % faster with the spare line, one or two
        // per cent for the step; two or more spare lines were slower)
        const int q = params->precision == FIPI_FP64 ? 8 : 16;  // complex values per 128 bytes
        ctx->d = Dims{params->nx, params->ny, params->nz, nxh, (nxh + q - 1) / q * q + q};
    }
    ctx->G = (size_t)params->nx * params->ny * params->nz;
    ctx->GH = (size_t)ctx->d.nxp * params->ny * params->nz;
    ctx->d.rs = ctx->G;
    ctx->d.hs = ctx->GH;
    ctx->n = params->switch_particle ? n_total : 0;
    hipError_t e = hipGetDevice(&ctx->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->pstream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) {
        rc = set_err(nullptr, FIPI_ERR_HIP, std::string("HIP init: ") + hipGetErrorString(e));
        delete ctx;
        return rc;
    }
    rc = ctx->fp64 ? setup_buffers<double>(ctx) : setup_buffers<float>(ctx);
    if (rc != FIPI_OK) {
        fipi_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return FIPI_OK;
}

extern "C" {

int fipi_create(const fipi_params *params, fipi_ctx **out)
{
    const int rc = create_ctx(params, 1, params ? params->n_particles : 0, out);
    // u stays spectral between steps unless the particle stage reads it
    if (rc == FIPI_OK) (*out)->u_lazy = params->interp_u == 0;
    return rc;
}

int fipi_destroy(fipi_ctx *ctx)
{
    if (!ctx) return FIPI_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->pstream) (void)hipStreamSynchronize(ctx->pstream);
    for (void *p : ctx->allocs) (void)hipFree(p);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->h_flags) (void)hipHostFree(ctx->h_flags);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->pstream) (void)hipStreamDestroy(ctx->pstream);
    delete ctx;
    return FIPI_OK;
}

const char *fipi_last_error(const fipi_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_last_error.c_str();
}

void *fipi_stream(fipi_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

}  // extern "C"

// State of the whole slab: phi[nbox*G] (box-major), n particles (pos[3n])
// with their boxes (box may be null for one box).
static int set_state_slab(fipi_ctx *ctx, const double *phi, const double *pos, const double *radius,
                          const double *cos_theta, const int *box)
{
    if (!ctx || !phi) return set_err(ctx, FIPI_ERR_ARG, "ctx and phi are required");
    if (ctx->n > 0 && !pos) return set_err(ctx, FIPI_ERR_ARG, "pos is required when particles are on");
    if (ctx->n > 0 && ctx->nbox > 1 && !box) return set_err(ctx, FIPI_ERR_ARG, "particle boxes are required");
    const size_t GS = ctx->G * ctx->nbox, GHS = ctx->GH * ctx->nbox;
    int rc;
    if ((rc = copy_in_real(ctx, ctx->c, phi, GS))) return rc;
    for (int a = 0; a < 3; a++) HIPCHK(ctx, hipMemsetAsync(ctx->u[a], 0, GS * ctx->esz, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(ctx->hist, 0, GHS * 2 * ctx->esz, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(ctx->d_flags, 0, sizeof(int) * 4, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(ctx->t_dev, 0, sizeof(int), ctx->stream));
    ctx->t = 0;
    ctx->cur = 0;
    ctx->conv_ready = false;
    ctx->u_stale = false;
    ctx->parts_stale = false;
    ctx->spread_ready = false;
    ctx->tables_ready = false;
    {
        // FP64 k = 0 mode of each box: c_hat(0) = sum of phi (used by the FP32-mixed path)
        std::vector<double> k0h((size_t)K0_STRIDE * ctx->nbox, 0.0);
        for (int b = 0; b < ctx->nbox; b++) {
            long double s = 0.0L;
            const double *pb = phi + (size_t)b * ctx->G;
            for (size_t i = 0; i < ctx->G; i++) s += (long double)pb[i];
            k0h[(size_t)K0_STRIDE * b] = (double)s;
        }
        HIPCHK(ctx, hipMemcpyAsync(ctx->k0, k0h.data(), sizeof(double) * k0h.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
        HIPCHK(ctx, hipMemsetAsync(ctx->conv0, 0, sizeof(double) * (ctx->nbox + 1), ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (ctx->fp64) Impl<double>::init_fields(ctx);
    else Impl<float>::init_fields(ctx);
    // keep phi exactly as given (the transform round trip above rewrote it)
    if ((rc = copy_in_real(ctx, ctx->c, phi, GS))) return rc;
    const int n = ctx->n;
    if (n > 0 && ctx->ps.box) {
        for (int i = 0; i < n; i++)
            if (box[i] < 0 || box[i] >= ctx->nbox) return set_err(ctx, FIPI_ERR_ARG, "particle box out of range");
        HIPCHK(ctx, hipMemcpyAsync((void *)ctx->ps.box, box, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    }
    if (n > 0) {
        PartState &ps = ctx->ps;
        double *tmp;
        HIPCHK(ctx, hipMalloc((void **)&tmp, sizeof(double) * 3 * n));
        HIPCHK(ctx, hipMemcpyAsync(tmp, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_aos_to_soa, grid1(n, 256), dim3(256), 0, ctx->stream, tmp, ps.p[0], ps.p[1], ps.p[2],
                           n);
        hipLaunchKernelGGL(k_aos_to_soa, grid1(n, 256), dim3(256), 0, ctx->stream, tmp, ps.po[0], ps.po[1],
                           ps.po[2], n);
        for (int a = 0; a < 3; a++) {
            HIPCHK(ctx, hipMemsetAsync(ps.u[a], 0, sizeof(double) * n, ctx->stream));
            HIPCHK(ctx, hipMemsetAsync(ps.uo[a], 0, sizeof(double) * n, ctx->stream));
            HIPCHK(ctx, hipMemsetAsync(ps.fpp[a], 0, sizeof(double) * n, ctx->stream));
            HIPCHK(ctx, hipMemsetAsync(ps.fpi[a], 0, sizeof(double) * n, ctx->stream));
            HIPCHK(ctx, hipMemsetAsync(ps.dcdx[a], 0, sizeof(double) * n, ctx->stream));
            HIPCHK(ctx, hipMemsetAsync(ps.uf[a], 0, sizeof(double) * n, ctx->stream));
            HIPCHK(ctx, hipMemsetAsync(ps.frand[a], 0, sizeof(double) * n, ctx->stream));
        }
        HIPCHK(ctx, hipMemsetAsync(ps.tau_i, 0, sizeof(double) * n, ctx->stream));
        HIPCHK(ctx, hipMemsetAsync(ps.mark, 0xff, sizeof(int), ctx->stream));
        std::vector<double> h(n);
        for (int i = 0; i < n; i++) h[i] = radius ? radius[i] : ctx->p.r_p;
        HIPCHK(ctx, hipMemcpyAsync(ps.rad, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < n; i++) h[i] = 1.0;  // element.cpp:34
        HIPCHK(ctx, hipMemcpyAsync(ps.cloc, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        ctx->pp.has_ctheta = 0;
        if (cos_theta) {
            for (int i = 0; i < n; i++) {
                h[i] = cos_theta[i];
                if (h[i] != 0.0) ctx->pp.has_ctheta = 1;
            }
        } else {
            for (int i = 0; i < n; i++) h[i] = 0.0;
        }
        HIPCHK(ctx, hipMemcpyAsync(ps.cth, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        HIPCHK(ctx, hipFree(tmp));
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    ctx->has_state = true;
    return FIPI_OK;
}

// set_state_slab for one box with device inputs (phi: G doubles; pos: 3n
// AoS; radius / cos_theta: n doubles or NULL).  The state equals what
// set_state_slab makes from the same values on the host.
__global__ void k_fill_double(double *a, int n, double v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_any_nonzero(const double *a, int n, int *flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && a[i] != 0.0) atomicOr(flag, 1);
}

static int copy_in_real_dev(fipi_ctx *ctx, void *dst, const double *d_src, size_t n)
{
    if (ctx->fp64) {
        HIPCHK(ctx, hipMemcpyAsync(dst, d_src, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        return FIPI_OK;
    }
    hipLaunchKernelGGL((k_from_double<float>), grid1(n, 256), dim3(256), 0, ctx->stream, d_src, (float *)dst, n);
    HIPCHK(ctx, hipGetLastError());
    return FIPI_OK;
}

static int set_state_dev(fipi_ctx *ctx, const double *d_phi, const double *d_pos, const double *d_radius,
                         const double *d_cth)
{
    const size_t G = ctx->G;
    int rc;
    if ((rc = copy_in_real_dev(ctx, ctx->c, d_phi, G))) return rc;
    for (int a = 0; a < 3; a++) HIPCHK(ctx, hipMemsetAsync(ctx->u[a], 0, G * ctx->esz, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(ctx->hist, 0, ctx->GH * 2 * ctx->esz, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(ctx->d_flags, 0, sizeof(int) * 4, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(ctx->t_dev, 0, sizeof(int), ctx->stream));
    ctx->t = 0;
    ctx->cur = 0;
    ctx->conv_ready = false;
    ctx->u_stale = false;
    ctx->parts_stale = false;
    ctx->spread_ready = false;
    ctx->tables_ready = false;
    {
        // the FP64 k = 0 mode (read by the FP32-mixed path): the same long
        // double host sum as set_state_slab, from one copy of phi
        std::vector<double> k0h((size_t)K0_STRIDE, 0.0);
        std::vector<double> h(G);
        HIPCHK(ctx, hipMemcpyAsync(h.data(), d_phi, G * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        long double sum = 0.0L;
        for (size_t i = 0; i < G; i++) sum += (long double)h[i];
        k0h[0] = (double)sum;
        HIPCHK(ctx, hipMemcpyAsync(ctx->k0, k0h.data(), sizeof(double) * k0h.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
        HIPCHK(ctx, hipMemsetAsync(ctx->conv0, 0, sizeof(double) * 2, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (ctx->fp64) Impl<double>::init_fields(ctx);
    else Impl<float>::init_fields(ctx);
    if ((rc = copy_in_real_dev(ctx, ctx->c, d_phi, G))) return rc;
    const int n = ctx->n;
    if (n > 0) {
        PartState &ps = ctx->ps;
        hipStream_t st = ctx->stream;
        hipLaunchKernelGGL(k_aos_to_soa, grid1(n, 256), dim3(256), 0, st, d_pos, ps.p[0], ps.p[1], ps.p[2], n);
        hipLaunchKernelGGL(k_aos_to_soa, grid1(n, 256), dim3(256), 0, st, d_pos, ps.po[0], ps.po[1], ps.po[2], n);
        for (int a = 0; a < 3; a++)
            for (double *z : {ps.u[a], ps.uo[a], ps.fpp[a], ps.fpi[a], ps.dcdx[a], ps.uf[a], ps.frand[a]})
                HIPCHK(ctx, hipMemsetAsync(z, 0, sizeof(double) * n, st));
        HIPCHK(ctx, hipMemsetAsync(ps.tau_i, 0, sizeof(double) * n, st));
        HIPCHK(ctx, hipMemsetAsync(ps.mark, 0xff, sizeof(int), st));
        if (d_radius) HIPCHK(ctx, hipMemcpyAsync(ps.rad, d_radius, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        else hipLaunchKernelGGL(k_fill_double, grid1(n, 256), dim3(256), 0, st, ps.rad, n, ctx->p.r_p);
        hipLaunchKernelGGL(k_fill_double, grid1(n, 256), dim3(256), 0, st, ps.cloc, n, 1.0);  // element.cpp:34
        ctx->pp.has_ctheta = 0;
        if (d_cth) {
            HIPCHK(ctx, hipMemcpyAsync(ps.cth, d_cth, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
            HIPCHK(ctx, hipMemsetAsync(ctx->d_flags + 3, 0, sizeof(int), st));
            hipLaunchKernelGGL(k_any_nonzero, grid1(n, 256), dim3(256), 0, st, (const double *)ps.cth, n,
                               ctx->d_flags + 3);
            int any = 0;
            HIPCHK(ctx, hipMemcpyAsync(&any, ctx->d_flags + 3, sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHK(ctx, hipStreamSynchronize(st));
            HIPCHK(ctx, hipMemsetAsync(ctx->d_flags + 3, 0, sizeof(int), st));
            ctx->pp.has_ctheta = any ? 1 : 0;
        } else {
            hipLaunchKernelGGL(k_fill_double, grid1(n, 256), dim3(256), 0, st, ps.cth, n, 0.0);
        }
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    ctx->has_state = true;
    return FIPI_OK;
}

extern "C" {

int fipi_set_state(fipi_ctx *ctx, const double *phi, const double *pos, const double *radius,
                   const double *cos_theta)
{
    return set_state_slab(ctx, phi, pos, radius, cos_theta, nullptr);
}

int fipi_set_state_device(fipi_ctx *ctx, const double *d_phi, const double *d_pos, const double *d_radius,
                          const double *d_cos_theta)
{
    if (!ctx || !d_phi) return set_err(ctx, FIPI_ERR_ARG, "ctx and phi are required");
    if (ctx->nbox != 1) return set_err(ctx, FIPI_ERR_UNSUPPORTED, "single-box contexts only");
    if (ctx->n > 0 && !d_pos) return set_err(ctx, FIPI_ERR_ARG, "pos is required when particles are on");
    return set_state_dev(ctx, d_phi, d_pos, d_radius, d_cos_theta);
}

int fipi_set_velocity(fipi_ctx *ctx, const double *u)
{
    if (!ctx || !u) return set_err(ctx, FIPI_ERR_ARG, "ctx and u are required");
    ctx->conv_ready = false;
    ctx->u_stale = false;
    for (int a = 0; a < 3; a++) {
        int rc = copy_in_real(ctx, ctx->u[a], u + (size_t)a * ctx->G, ctx->G);
        if (rc) return rc;
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return FIPI_OK;
}

int fipi_get_phi(fipi_ctx *ctx, double *phi)
{
    if (!ctx || !phi) return set_err(ctx, FIPI_ERR_ARG, "ctx and phi are required");
    return copy_out_real(ctx, phi, ctx->c, ctx->G);
}

int fipi_get_velocity(fipi_ctx *ctx, double *u)
{
    if (!ctx || !u) return set_err(ctx, FIPI_ERR_ARG, "ctx and u are required");
    ensure_real_u(ctx);
    for (int a = 0; a < 3; a++) {
        int rc = copy_out_real(ctx, u + (size_t)a * ctx->G, ctx->u[a], ctx->G);
        if (rc) return rc;
    }
    return FIPI_OK;
}

int fipi_get_grad_phi(fipi_ctx *ctx, double *g)
{
    if (!ctx || !g) return set_err(ctx, FIPI_ERR_ARG, "ctx and grad are required");
    for (int a = 0; a < 3; a++) {
        int rc = copy_out_real(ctx, g + (size_t)a * ctx->G, ctx->g[a], ctx->G);
        if (rc) return rc;
    }
    return FIPI_OK;
}

int fipi_get_particles(fipi_ctx *ctx, double *pos, double *pos_old, double *u, double *u_old, double *f_pp,
                       double *f_pi)
{
    if (!ctx) return set_err(ctx, FIPI_ERR_ARG, "ctx is NULL");
    const int n = ctx->n;
    if (n == 0) return FIPI_OK;
    ensure_particles(ctx);
    double *tmp;
    HIPCHK(ctx, hipMalloc((void **)&tmp, sizeof(double) * 3 * n));
    double *const *src[6] = {ctx->ps.p, ctx->ps.po, ctx->ps.u, ctx->ps.uo, ctx->ps.fpp, ctx->ps.fpi};
    double *dst[6] = {pos, pos_old, u, u_old, f_pp, f_pi};
    for (int k = 0; k < 6; k++) {
        if (!dst[k]) continue;
        hipLaunchKernelGGL(k_soa_to_aos, grid1(n, 256), dim3(256), 0, ctx->stream, src[k][0], src[k][1], src[k][2],
                           tmp, n);
        HIPCHK(ctx, hipMemcpyAsync(dst[k], tmp, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    HIPCHK(ctx, hipFree(tmp));
    return FIPI_OK;
}

int fipi_get_time(fipi_ctx *ctx, int32_t *t)
{
    if (!ctx || !t) return set_err(ctx, FIPI_ERR_ARG, "ctx and t are required");
    *t = ctx->t;
    return FIPI_OK;
}

}  // extern "C"

// Diagnostics of box b; its particles are [p0, p1) of the slab's list.
static int diag_box(fipi_ctx *ctx, int b, int p0, int p1, fipi_diag *out)
{
    if (!ctx || !out) return set_err(ctx, FIPI_ERR_ARG, "ctx and out are required");
    if (b < 0 || b >= ctx->nbox) return set_err(ctx, FIPI_ERR_ARG, "box out of range");
    ensure_particles(ctx);
    hipStream_t st = ctx->stream;
    const double dx = ctx->p.lx / ctx->p.nx;
    const size_t off = (size_t)b * ctx->G * ctx->esz;
    auto at = [&](void *p) { return (const void *)((const char *)p + off); };
    if (ctx->fp64)
        hipLaunchKernelGGL((k_diag_partial<double>), dim3(DIAG_BLOCKS), dim3(256), 0, st, (const double *)at(ctx->c),
                           (const double *)at(ctx->g[0]), (const double *)at(ctx->g[1]),
                           (const double *)at(ctx->g[2]), ctx->G, ctx->p.lambda, ctx->p.epsilon, ctx->d_red);
    else
        hipLaunchKernelGGL((k_diag_partial<float>), dim3(DIAG_BLOCKS), dim3(256), 0, st, (const float *)at(ctx->c),
                           (const float *)at(ctx->g[0]), (const float *)at(ctx->g[1]), (const float *)at(ctx->g[2]),
                           ctx->G, ctx->p.lambda, ctx->p.epsilon, ctx->d_red);
    hipLaunchKernelGGL(k_sum_strided, dim3(1), dim3(256), 0, st, ctx->d_red, DIAG_BLOCKS, 3, 3,
                       ctx->d_red + 3 * DIAG_BLOCKS);
    const int np = p1 - p0;
    if (np > 0)
        hipLaunchKernelGGL(k_sum_strided, dim3(1), dim3(256), 0, st, ctx->ps.tau_i + p0, np, 1, 1,
                           ctx->d_red + 3 * DIAG_BLOCKS + 3);
    double h[4] = {0, 0, 0, 0};
    HIPCHK(ctx, hipMemcpyAsync(h, ctx->d_red + 3 * DIAG_BLOCKS, sizeof(double) * (np > 0 ? 4 : 3),
                               hipMemcpyDeviceToHost, st));
    int mark = -1;
    if (ctx->n > 0) HIPCHK(ctx, hipMemcpyAsync(&mark, ctx->ps.mark, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    const double cell = dx * dx * dx;  // cubic(dx), field.cpp:512
    out->t = ctx->t;
    out->it_mark = mark;
    out->mean_phi = h[0] / (double)ctx->G;
    if (!ctx->fp64) {
        // FP32-mixed: the mass is carried in FP64 as the k = 0 mode
        double k0h[3];
        HIPCHK(ctx, hipMemcpy(k0h, ctx->k0 + (size_t)K0_STRIDE * b, sizeof(k0h), hipMemcpyDeviceToHost));
        out->mean_phi = k0h[ctx->cur] / (double)ctx->G;
    }
    out->free_energy = h[1] * cell;
    out->interface_area = h[2] * cell;
    // the tau terms count only when it_mark's attachment force is non-zero
    // (cluster.cpp:256-259); the step's stage 2 records that gate
    int mark_on = 0;
    if (ctx->n > 0) HIPCHK(ctx, hipMemcpy(&mark_on, ctx->ps.mark_on, sizeof(int), hipMemcpyDeviceToHost));
    out->tau = mark_on ? h[3] : 0.0;
    return FIPI_OK;
}

extern "C" {

int fipi_diagnostics(fipi_ctx *ctx, fipi_diag *out)
{
    if (!ctx || !out) return set_err(ctx, FIPI_ERR_ARG, "ctx and out are required");
    return diag_box(ctx, 0, 0, ctx->n, out);
}

int fipi_step_async(fipi_ctx *ctx, int32_t nsteps)
{
    if (!ctx) return set_err(ctx, FIPI_ERR_ARG, "ctx is NULL");
    if (!ctx->has_state) return set_err(ctx, FIPI_ERR_STATE, "fipi_set_state must be called first");
    if (nsteps < 0) return set_err(ctx, FIPI_ERR_ARG, "nsteps < 0");
    // each step's particle stage runs on a second stream beside the next
    // step's first passes (never under per-launch timing, whose events are
    // on the main stream); the main stream waits for the last one, so all
    // work ends there
    const bool beside = ctx->c_alt != nullptr && !ctx->prof_on && !ctx->fix_particles;
    for (int s = 0; s < nsteps; s++) {
        const bool last = s + 1 == nsteps;
        if (ctx->fp64) Impl<double>::step(ctx, beside, last);
        else Impl<float>::step(ctx, beside, last);
    }
    HIPCHK(ctx, hipGetLastError());
    return FIPI_OK;
}

int fipi_sync(fipi_ctx *ctx)
{
    if (!ctx) return set_err(ctx, FIPI_ERR_ARG, "ctx is NULL");
    return check_flags(ctx);
}

int fipi_step(fipi_ctx *ctx, int32_t nsteps)
{
    int rc = fipi_step_async(ctx, nsteps);
    if (rc) return rc;
    return check_flags(ctx);
}

int fipi_profile_steps(fipi_ctx *ctx, int32_t nsteps, fipi_kernel_stat *out, int32_t max_out, int32_t *count)
{
    if (!ctx || !count || max_out < 0 || (max_out > 0 && !out)) return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->prof.clear();
    ctx->ev_used = 0;
    ctx->prof_on = true;
    int rc = fipi_step_async(ctx, nsteps);
    ctx->prof_on = false;
    if (rc) return rc;
    if ((rc = check_flags(ctx))) return rc;
    std::vector<fipi_kernel_stat> sites;
    for (const auto &r : ctx->prof) {
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, r.a, r.b));
        size_t k = 0;
        while (k < sites.size() && std::strncmp(sites[k].name, r.name, sizeof(sites[k].name)) != 0) k++;
        if (k == sites.size()) {
            fipi_kernel_stat s{};
            std::strncpy(s.name, r.name, sizeof(s.name) - 1);
            sites.push_back(s);
        }
        sites[k].launches++;
        sites[k].total_ms += ms;
        sites[k].bytes += r.bytes;
    }
    *count = (int32_t)sites.size();
    for (size_t k = 0; k < sites.size() && (int32_t)k < max_out; k++) out[k] = sites[k];
    ctx->prof.clear();
    return FIPI_OK;
}

int fipi_fft_r2c(fipi_ctx *ctx, const void *d_real, void *d_half)
{
    if (!ctx || !d_real || !d_half) return set_err(ctx, FIPI_ERR_ARG, "null argument");
    if (ctx->fp64) Impl<double>::r2c(ctx, (const double *)d_real, (double2 *)d_half);
    else Impl<float>::r2c(ctx, (const float *)d_real, (float2 *)d_half);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    return FIPI_OK;
}

int fipi_fft_c2r(fipi_ctx *ctx, void *d_half, void *d_real)
{
    if (!ctx || !d_real || !d_half) return set_err(ctx, FIPI_ERR_ARG, "null argument");
    const double s = 1.0 / (double)ctx->G;
    if (ctx->fp64) {
        Impl<double>::c2r(ctx, (double2 *)d_half, (double *)d_real);
        hipLaunchKernelGGL((k_scale<double>), grid1(ctx->G, 256), dim3(256), 0, ctx->stream, (double *)d_real,
                           ctx->G, s);
    } else {
        Impl<float>::c2r(ctx, (float2 *)d_half, (float *)d_real);
        hipLaunchKernelGGL((k_scale<float>), grid1(ctx->G, 256), dim3(256), 0, ctx->stream, (float *)d_real,
                           ctx->G, s);
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    return FIPI_OK;
}

int fipi_cells_build(fipi_ctx *ctx, const double *d_pos, int32_t n, int32_t *d_cell_start, int32_t *d_ids,
                     int32_t *ncell_out)
{
    if (!ctx || (n > 0 && (!d_pos || !d_ids)) || !d_cell_start || n < 0)
        return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    HIPCHK(ctx, hipMemsetAsync(ctx->d_flags, 0, sizeof(int) * 4, ctx->stream));
    TmpParts tp;
    int rc = tp.init(ctx, d_pos, n);
    if (rc) return rc;
    if ((rc = unit_cells(ctx, tp, d_cell_start, d_ids))) return rc;
    if (ncell_out) *ncell_out = ctx->ncell;
    return check_flags(ctx);
}

int fipi_contact_pairs(fipi_ctx *ctx, const double *d_pos, int32_t n, int32_t *d_pairs, int64_t max_pairs,
                       int64_t *count)
{
    if (!ctx || n < 0 || (n > 0 && !d_pos) || !count || (max_pairs > 0 && !d_pairs))
        return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    HIPCHK(ctx, hipMemsetAsync(ctx->d_flags, 0, sizeof(int) * 4, ctx->stream));
    TmpParts tp;
    int rc = tp.init(ctx, d_pos, n);
    if (rc) return rc;
    int *ids, *pcnt, *pstart;
    if ((rc = tp.alloc((void **)&ids, sizeof(int) * (n > 0 ? n : 1)))) return rc;
    if ((rc = tp.alloc((void **)&pcnt, sizeof(int) * (n + 1)))) return rc;
    if ((rc = tp.alloc((void **)&pstart, sizeof(int) * (n + 1)))) return rc;
    if ((rc = unit_cells(ctx, tp, ctx->cell_start, ids))) return rc;
    PartParams pp = ctx->pp;
    pp.n = n;
    hipStream_t st = ctx->stream;
    if (n > 0) {
        hipLaunchKernelGGL(k_pairs, grid1(n, 128), dim3(128), 0, st, pp, tp.ps, ctx->cell_start, ids,
                           (const int *)nullptr, pcnt, (int *)nullptr, (int64_t)0);
        int *bs;
        if ((rc = tp.alloc((void **)&bs, sizeof(int) * (n / SCAN_ITEMS + 2)))) return rc;
        scan(ctx, pcnt, pstart, n, bs, nullptr, nullptr, 0, nullptr);
        if (max_pairs > 0) {
            hipLaunchKernelGGL(k_pairs, grid1(n, 128), dim3(128), 0, st, pp, tp.ps, ctx->cell_start, ids,
                               (const int *)pstart, pcnt, d_pairs, max_pairs);
            hipLaunchKernelGGL(k_pairs_sort, grid1(n, 256), dim3(256), 0, st, n, pstart, d_pairs, max_pairs);
        }
        int tot = 0;
        HIPCHK(ctx, hipMemcpyAsync(&tot, pstart + n, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        *count = tot;
    } else {
        *count = 0;
    }
    return check_flags(ctx);
}

int fipi_contact_forces(fipi_ctx *ctx, const double *d_pos, int32_t n, int32_t t, double *d_f_pp)
{
    if (!ctx || n < 0 || (n > 0 && (!d_pos || !d_f_pp))) return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    HIPCHK(ctx, hipMemsetAsync(ctx->d_flags, 0, sizeof(int) * 4, ctx->stream));
    TmpParts tp;
    int rc = tp.init(ctx, d_pos, n);
    if (rc) return rc;
    int *ids;
    if ((rc = tp.alloc((void **)&ids, sizeof(int) * (n > 0 ? n : 1)))) return rc;
    if ((rc = unit_cells(ctx, tp, ctx->cell_start, ids))) return rc;
    PartParams pp = ctx->pp;
    pp.n = n;
    const double thr = ctx->p.ramp_t;
    const double rmp = t < 0 ? 0.0 : (t < thr ? t / thr : 1.0);
    if (n > 0) {
        launch_contact(ctx->stream, n, pp, tp.ps, ctx->cell_start, ids, rmp, (const int *)nullptr, 0.0,
                       ctx->d_flags + 1);
        hipLaunchKernelGGL(k_soa_to_aos, grid1(n, 256), dim3(256), 0, ctx->stream, tp.ps.fpp[0], tp.ps.fpp[1],
                           tp.ps.fpp[2], d_f_pp, n);
    }
    return check_flags(ctx);
}

int fipi_interp(fipi_ctx *ctx, const void *d_grid, const double *d_pos, int32_t n, double *d_out)
{
    if (!ctx || !d_grid || n < 0 || (n > 0 && (!d_pos || !d_out))) return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    TmpParts tp;
    int rc = tp.init(ctx, d_pos, n);
    if (rc) return rc;
    PartParams pp = ctx->pp;
    pp.n = n;
    if (n > 0) {
        if (ctx->fp64)
            hipLaunchKernelGGL((k_interp_field<double>), grid1(n, 256), dim3(256), 0, ctx->stream, pp,
                               (const double *)d_grid, tp.ps.p[0], tp.ps.p[1], tp.ps.p[2], d_out);
        else
            hipLaunchKernelGGL((k_interp_field<float>), grid1(n, 256), dim3(256), 0, ctx->stream, pp,
                               (const float *)d_grid, tp.ps.p[0], tp.ps.p[1], tp.ps.p[2], d_out);
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    return FIPI_OK;
}

int fipi_spread(fipi_ctx *ctx, const double *d_pos, const double *d_force, int32_t n, void *d_grid3)
{
    if (!ctx || !d_grid3 || n < 0 || (n > 0 && (!d_pos || !d_force)))
        return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    if (n > ctx->n && n > 0) {
        // entry buffer is sized for the context's particle count
        if (ctx->n < n) return set_err(ctx, FIPI_ERR_ARG, "n exceeds the context's n_particles");
    }
    HIPCHK(ctx, hipMemsetAsync(ctx->d_flags, 0, sizeof(int) * 4, ctx->stream));
    TmpParts tp;
    int rc = tp.init(ctx, d_pos, n);
    if (rc) return rc;
    double *f[3];
    for (int a = 0; a < 3; a++)
        if ((rc = tp.alloc((void **)&f[a], sizeof(double) * (n > 0 ? n : 1)))) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_aos_to_soa, grid1(n, 256), dim3(256), 0, ctx->stream, d_force, f[0], f[1], f[2], n);
    PartParams pp = ctx->pp;
    pp.n = n;
    HIPCHK(ctx, hipMemsetAsync(d_grid3, 0, 3 * ctx->G * ctx->esz, ctx->stream));
    if (ctx->fp64) {
        Impl<double>::spread_build(ctx, pp, tp.ps.p[0], tp.ps.p[1], tp.ps.p[2], f[0], f[1], f[2], nullptr);
        hipLaunchKernelGGL((k_spread_dense<double>), grid1(ctx->nlines, 256), dim3(256), 0, ctx->stream, ctx->nlines,
                           ctx->p.nx, ctx->G, ctx->line_start, (const int *)ctx->node_cnt,
                           (const SpreadNode<double> *)ctx->spr_tmp, (double *)d_grid3);
    } else {
        Impl<float>::spread_build(ctx, pp, tp.ps.p[0], tp.ps.p[1], tp.ps.p[2], f[0], f[1], f[2], nullptr);
        hipLaunchKernelGGL((k_spread_dense<float>), grid1(ctx->nlines, 256), dim3(256), 0, ctx->stream, ctx->nlines,
                           ctx->p.nx, ctx->G, ctx->line_start, (const int *)ctx->node_cnt,
                           (const SpreadNode<float> *)ctx->spr_tmp, (float *)d_grid3);
    }
    return check_flags(ctx);
}

int fipi_interp2(fipi_ctx *ctx, const void *d_grid, const double *d_pos, int32_t n, double *d_out)
{
    if (!ctx || !d_grid || n < 0 || (n > 0 && (!d_pos || !d_out))) return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    TmpParts tp;
    int rc = tp.init(ctx, d_pos, n);
    if (rc) return rc;
    PartParams pp = ctx->pp;
    pp.n = n;
    if (n > 0) {
        if (ctx->fp64)
            hipLaunchKernelGGL((k_interp2_field<double>), grid1(n, 256), dim3(256), 0, ctx->stream, pp,
                               (const double *)d_grid, tp.ps.p[0], tp.ps.p[1], tp.ps.p[2], d_out);
        else
            hipLaunchKernelGGL((k_interp2_field<float>), grid1(n, 256), dim3(256), 0, ctx->stream, pp,
                               (const float *)d_grid, tp.ps.p[0], tp.ps.p[1], tp.ps.p[2], d_out);
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    return FIPI_OK;
}

int fipi_brownian_forces(fipi_ctx *ctx, int32_t t, int32_t stage, int32_t n, double *d_out)
{
    if (!ctx || n < 0 || (n > 0 && !d_out) || (stage != 1 && stage != 2) || t < 0)
        return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    if (n == 0) return FIPI_OK;
    PartParams pp = ctx->pp;
    pp.n = n;
    pp.brown_amp = ctx->p.kt > 0.0 ? std::sqrt(12.0 * ctx->p.pi * ctx->p.r_p * ctx->p.mu * ctx->p.kt / ctx->p.dt) : 0.0;
    hipLaunchKernelGGL(k_brownian_unit, grid1(n, 256), dim3(256), 0, ctx->stream, pp, (int)t, (int)stage, d_out);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    return FIPI_OK;
}

int fipi_get_particle_fluid(fipi_ctx *ctx, double *uf, double *f_rand, double *f_drag)
{
    if (!ctx) return set_err(ctx, FIPI_ERR_ARG, "ctx is NULL");
    if (ctx->nbox != 1) return set_err(ctx, FIPI_ERR_UNSUPPORTED, "single-box contexts only");
    const int n = ctx->n;
    if (n == 0) return FIPI_OK;
    if (ctx->p_pending) HIPCHK(ctx, hipStreamSynchronize(ctx->pstream));
    ensure_particles(ctx);
    const PartState &ps = ctx->ps;
    double *tmp;
    HIPCHK(ctx, hipMalloc((void **)&tmp, sizeof(double) * 3 * n));
    double *out[3] = {uf, f_rand, f_drag};
    int rc = FIPI_OK;
    for (int w = 0; w < 3 && rc == FIPI_OK; w++) {
        if (!out[w]) continue;
        if (w < 2) {
            double *const *src = w == 0 ? ps.uf : ps.frand;
            hipLaunchKernelGGL(k_soa_to_aos, grid1(n, 256), dim3(256), 0, ctx->stream, (const double *)src[0],
                               (const double *)src[1], (const double *)src[2], tmp, n);
        } else {
            hipLaunchKernelGGL(k_drag_unit, grid1(n, 256), dim3(256), 0, ctx->stream, ctx->pp, ps, tmp, n);
        }
        hipError_t e = hipMemcpyAsync(out[w], tmp, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = set_err(ctx, FIPI_ERR_HIP, std::string("particle fluid copy: ") + hipGetErrorString(e));
    }
    (void)hipFree(tmp);
    return rc;
}

}  // extern "C"

namespace {

// tuning choices of a context (fipi_get_tuning / fipi_set_tuning)
void y_plans_of(const fipi_ctx *ctx, int *def, int *alt)
{
    if (ctx->fp64) Impl<double>::y_plans(ctx, def, alt);
    else Impl<float>::y_plans(ctx, def, alt);
}

void tuning_out(const fipi_ctx *ctx, fipi_tuning_t *out)
{
    int def, alt;
    y_plans_of(ctx, &def, &alt);
    out->y_tile_bytes = ctx->tune.y_tb;
    out->hub_tile_bytes = ctx->tune.hub_tb;
    out->stokes_tile_bytes = ctx->tune.stokes_tb;
    out->y_values_per_thread = ctx->tune.y_e != 0 ? ctx->tune.y_e : def;
}

int tuning_in(fipi_ctx *ctx, const fipi_tuning_t *in)
{
    auto tb_ok = [](int v) { return v == 64 || v == 128; };
    if (!tb_ok(in->y_tile_bytes) || !tb_ok(in->hub_tile_bytes) || !tb_ok(in->stokes_tile_bytes))
        return set_err(ctx, FIPI_ERR_ARG, "tile widths are 64 or 128 bytes");
    int def, alt;
    y_plans_of(ctx, &def, &alt);
    const int e = in->y_values_per_thread;
    if (e != 0 && e != def && (alt == 0 || e != alt))
        return set_err(ctx, FIPI_ERR_ARG, "no y-pass plan with that many values per thread for this grid");
    ctx->tune.y_tb = in->y_tile_bytes;
    ctx->tune.hub_tb = in->hub_tile_bytes;
    ctx->tune.stokes_tb = in->stokes_tile_bytes;
    ctx->tune.y_e = (e == 0 || e == def) ? 0 : e;
    return FIPI_OK;
}

}  // namespace

#include "ensemble.hpp"
#include "adjoint.hpp"
#include "facade_ops.hpp"
#include "init_device.hpp"

// ---------------------------------------------------------------------------
// Snapshot / restart: the whole state of a context (fields, SBDF2 history
// and ring index, mass mode, particles, step counters) as one host buffer
// that fipi_snapshot_load restores into a context created with the same
// parameters; the run then continues bit for bit.  The reference writes
// ASCII VTK output without the history (field.cpp:469-518,
// cluster.cpp:415-456), which cannot restart a run.
// ---------------------------------------------------------------------------
namespace {

constexpr int32_t SNAP_VERSION = 2;

struct SnapHeader {
    char magic[8];
    // y_e: the y-pass FFT plan in use (values per thread, 0 = the default);
    // it changes rounding, so a restore takes it over
    int32_t version, nx, ny, nz, n, nbox, precision, t, cur, conv_ready, has_ctheta, y_e;
    int64_t bytes;
};

std::vector<std::pair<void *, size_t>> snapshot_buffers(fipi_ctx *ctx)
{
    const size_t nb = ctx->nbox;
    const size_t R = ctx->G * nb * ctx->esz, H = ctx->GH * nb * 2 * ctx->esz;
    std::vector<std::pair<void *, size_t>> v = {
        {ctx->c, R},          {ctx->g[0], R},  {ctx->g[1], R},  {ctx->g[2], R},  {ctx->u[0], R},
        {ctx->u[1], R},       {ctx->u[2], R},  {ctx->ck[0], H}, {ctx->ck[1], H}, {ctx->fk, H},
        {ctx->hist, H},       {ctx->h[0], H},  {ctx->k0, sizeof(double) * K0_STRIDE * nb},
        {ctx->conv0, sizeof(double) * (nb + 1)},
        {ctx->conv_part, sizeof(double) * nb * (size_t)(ctx->fp64 ? xblocks<double>(ctx->d) : xblocks<float>(ctx->d))},
        {ctx->t_dev, sizeof(int)}};
    const size_t n = ctx->n;
    if (n > 0) {
        PartState &ps = ctx->ps;
        const size_t D = sizeof(double) * n;
        for (int a = 0; a < 3; a++) {
            v.push_back({ps.p[a], D});
            v.push_back({ps.po[a], D});
            v.push_back({ps.u[a], D});
            v.push_back({ps.uo[a], D});
            v.push_back({ps.fpp[a], D});
            v.push_back({ps.fpi[a], D});
            v.push_back({ps.dcdx[a], D});
            if (ctx->pp.interp_u) v.push_back({ps.uf[a], D});
            if (ctx->pp.brown_amp > 0.0) v.push_back({ps.frand[a], D});
        }
        v.push_back({ps.rad, D});
        v.push_back({ps.cth, D});
        v.push_back({ps.cloc, D});
        v.push_back({ps.tau_i, D});
        v.push_back({ps.mark, sizeof(int)});
        v.push_back({ps.mark_on, sizeof(int)});
        if (ps.box) v.push_back({(void *)ps.box, sizeof(int) * n});
    }
    return v;
}

}  // namespace

extern "C" {

int64_t fipi_snapshot_size(fipi_ctx *ctx)
{
    if (!ctx) return set_err(ctx, FIPI_ERR_ARG, "ctx is NULL");
    int64_t b = sizeof(SnapHeader);
    for (auto &x : snapshot_buffers(ctx)) b += (int64_t)x.second;
    return b;
}

int fipi_snapshot_save(fipi_ctx *ctx, void *buf, int64_t size)
{
    if (!ctx || !buf) return set_err(ctx, FIPI_ERR_ARG, "ctx and buf are required");
    if (!ctx->has_state) return set_err(ctx, FIPI_ERR_STATE, "no state to save");
    ensure_real_u(ctx);
    ensure_particles(ctx);
    const int64_t need = fipi_snapshot_size(ctx);
    if (size < need) return set_err(ctx, FIPI_ERR_OVERFLOW, "snapshot buffer too small");
    SnapHeader h{};
    std::memcpy(h.magic, "FIPISNAP", 8);
    h.version = SNAP_VERSION;
    h.nx = ctx->p.nx;
    h.ny = ctx->p.ny;
    h.nz = ctx->p.nz;
    h.n = ctx->n;
    h.nbox = ctx->nbox;
    h.precision = ctx->p.precision;
    h.t = ctx->t;
    h.cur = ctx->cur;
    h.conv_ready = ctx->conv_ready ? 1 : 0;
    h.has_ctheta = ctx->pp.has_ctheta;
    h.y_e = ctx->tune.y_e;
    h.bytes = need;
    std::memcpy(buf, &h, sizeof(h));
    char *o = (char *)buf + sizeof(h);
    for (auto &x : snapshot_buffers(ctx)) {
        HIPCHK(ctx, hipMemcpyAsync(o, x.first, x.second, hipMemcpyDeviceToHost, ctx->stream));
        o += x.second;
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return FIPI_OK;
}

int fipi_snapshot_load(fipi_ctx *ctx, const void *buf, int64_t size)
{
    if (!ctx || !buf) return set_err(ctx, FIPI_ERR_ARG, "ctx and buf are required");
    if (size < (int64_t)sizeof(SnapHeader)) return set_err(ctx, FIPI_ERR_ARG, "not a snapshot");
    SnapHeader h;
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, "FIPISNAP", 8) != 0 || h.version != SNAP_VERSION)
        return set_err(ctx, FIPI_ERR_ARG, "not a snapshot of this format version");
    if (h.nx != ctx->p.nx || h.ny != ctx->p.ny || h.nz != ctx->p.nz || h.n != ctx->n || h.nbox != ctx->nbox ||
        h.precision != ctx->p.precision)
        return set_err(ctx, FIPI_ERR_ARG, "snapshot of a different grid, particle count or precision");
    if (h.bytes != fipi_snapshot_size(ctx) || size < h.bytes) return set_err(ctx, FIPI_ERR_ARG, "snapshot size mismatch");
    {
        int def, alt;
        y_plans_of(ctx, &def, &alt);
        if (h.y_e != 0 && h.y_e != alt) return set_err(ctx, FIPI_ERR_ARG, "snapshot names an unknown y-pass plan");
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const char *in = (const char *)buf + sizeof(h);
    for (auto &x : snapshot_buffers(ctx)) {
        HIPCHK(ctx, hipMemcpyAsync(x.first, in, x.second, hipMemcpyHostToDevice, ctx->stream));
        in += x.second;
    }
    HIPCHK(ctx, hipMemsetAsync(ctx->d_flags, 0, sizeof(int) * 4, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->t = h.t;
    ctx->cur = h.cur;
    ctx->conv_ready = h.conv_ready != 0;
    ctx->u_stale = false;
    ctx->parts_stale = false;
    ctx->pp.has_ctheta = h.has_ctheta;
    ctx->tune.y_e = h.y_e;
    ctx->spread_ready = false;
    ctx->tables_ready = false;
    ctx->p_pending = false;
    ctx->has_state = true;
    return FIPI_OK;
}

int fipi_tune(fipi_ctx *ctx, int32_t steps)
{
    if (!ctx || steps < 1) return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    if (!ctx->has_state) return set_err(ctx, FIPI_ERR_STATE, "fipi_set_state must be called first");
    struct DevBuf {
        void *p = nullptr;
        hipEvent_t a = nullptr, b = nullptr;
        ~DevBuf()
        {
            if (p) (void)hipFree(p);
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } keep;
    ensure_real_u(ctx);
    ensure_particles(ctx);
    size_t total = 0;
    for (auto &x : snapshot_buffers(ctx)) total += x.second;
    HIPCHK(ctx, hipMalloc(&keep.p, total));
    HIPCHK(ctx, hipEventCreate(&keep.a));
    HIPCHK(ctx, hipEventCreate(&keep.b));
    {
        char *o = (char *)keep.p;
        for (auto &x : snapshot_buffers(ctx)) {
            HIPCHK(ctx, hipMemcpyAsync(o, x.first, x.second, hipMemcpyDeviceToDevice, ctx->stream));
            o += x.second;
        }
    }
    const int t0 = ctx->t, cur0 = ctx->cur;
    const bool conv0 = ctx->conv_ready;
    const fipi_ctx::Tune before = ctx->tune;
    fipi_ctx::Tune best = before;
    float best_ms = 1e30f;
    int ydef, yalt;
    y_plans_of(ctx, &ydef, &yalt);
    std::vector<int> plans = {0};  // 0: the default plan of the size
    if (yalt != 0) plans.push_back(yalt);
    // time every combination; the first error ends the trials
    auto trials = [&]() -> int {
        for (int ye : plans)
            for (int y : {64, 128})
                for (int h : {64, 128})
                    for (int st : {64, 128}) {
                        ctx->tune = fipi_ctx::Tune{y, h, st, ye};
                        int rc = fipi_step_async(ctx, 1);  // warm-up (first use of these kernels)
                        if (rc) return rc;
                        HIPCHK(ctx, hipEventRecord(keep.a, ctx->stream));
                        if ((rc = fipi_step_async(ctx, steps))) return rc;
                        HIPCHK(ctx, hipEventRecord(keep.b, ctx->stream));
                        HIPCHK(ctx, hipEventSynchronize(keep.b));
                        float ms = 0.f;
                        HIPCHK(ctx, hipEventElapsedTime(&ms, keep.a, keep.b));
                        if (ms < best_ms) {
                            best_ms = ms;
                            best = ctx->tune;
                        }
                    }
        return check_flags(ctx);
    };
    const int rc = trials();
    ctx->tune = rc == FIPI_OK ? best : before;
    // back to the state before the trials, also after a failed one (into
    // the arrays in use now: the spare c / grad c set may have been swapped in)
    (void)hipStreamSynchronize(ctx->stream);
    const char *in = (const char *)keep.p;
    for (auto &x : snapshot_buffers(ctx)) {
        HIPCHK(ctx, hipMemcpyAsync(x.first, in, x.second, hipMemcpyDeviceToDevice, ctx->stream));
        in += x.second;
    }
    HIPCHK(ctx, hipMemsetAsync(ctx->d_flags, 0, sizeof(int) * 4, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->t = t0;
    ctx->cur = cur0;
    ctx->conv_ready = conv0;
    ctx->u_stale = false;
    ctx->parts_stale = false;
    ctx->spread_ready = false;
    ctx->tables_ready = false;
    return rc;
}

int fipi_get_tuning(const fipi_ctx *ctx, fipi_tuning_t *out)
{
    if (!ctx || !out) return set_err(nullptr, FIPI_ERR_ARG, "ctx and out are required");
    tuning_out(ctx, out);
    return FIPI_OK;
}

int fipi_set_tuning(fipi_ctx *ctx, const fipi_tuning_t *in)
{
    if (!ctx || !in) return set_err(ctx, FIPI_ERR_ARG, "ctx and in are required");
    return tuning_in(ctx, in);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Cell list without allocation: fipi_cells_build_noalloc.  Every temporary
// of the build (positions in SoA, per-particle cells, cell counts and
// cursors, scan block totals, argmax-y partials, the error flag) lives in
// one static device array, `scratch`, of NOALLOC_SCRATCH_BYTES per GPU; its
// address on each device comes from hipGetSymbolAddress on the first call
// made with that device current.  The call allocates nothing and never
// synchronises: counters are cleared by a kernel at the start of each call
// (no memset node, so the call may be captured in a HIP graph).  Calls on
// one GPU share the array: two calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace {

constexpr int NA_MAX_N = FIPI_NOALLOC_MAX_PARTICLES;
constexpr int NA_MAX_CELLS = FIPI_NOALLOC_MAX_CELLS;
constexpr size_t NA_BSUM = (size_t)(NA_MAX_N > NA_MAX_CELLS ? NA_MAX_N : NA_MAX_CELLS) / SCAN_ITEMS + 2;
constexpr size_t NA_PART = (size_t)NA_MAX_N / 256 + 2;
// layout in bytes, every region 256-byte aligned
constexpr size_t na_up(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NA_OFF_P = 0;                                               // 3 x double[N]
constexpr size_t NA_OFF_CELL = NA_OFF_P + 3 * na_up(sizeof(double) * NA_MAX_N);  // int[N]
constexpr size_t NA_OFF_CNT = NA_OFF_CELL + na_up(sizeof(int) * NA_MAX_N);      // int[C+1]
constexpr size_t NA_OFF_CUR = NA_OFF_CNT + na_up(sizeof(int) * (NA_MAX_CELLS + 1));  // int[C+1]
constexpr size_t NA_OFF_BSUM = NA_OFF_CUR + na_up(sizeof(int) * (NA_MAX_CELLS + 1));
constexpr size_t NA_OFF_YMAX = NA_OFF_BSUM + na_up(sizeof(int) * NA_BSUM);
constexpr size_t NA_OFF_YIDX = NA_OFF_YMAX + na_up(sizeof(double) * NA_PART);
constexpr size_t NA_OFF_MISC = NA_OFF_YIDX + na_up(sizeof(int) * NA_PART);  // mark, error flags
constexpr size_t NOALLOC_SCRATCH_BYTES = NA_OFF_MISC + 256;
static_assert(NOALLOC_SCRATCH_BYTES < (size_t)16 << 20, "the no-allocation scratch stays below 16 MiB");

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mutex;
char *g_na_ptr[NA_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (resolved once per
// device); nullptr and a message on failure.
char *na_scratch(int *dev_out, std::string *err)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return nullptr;
    }
    if (dev < 0 || dev >= NA_MAX_DEVICES) {
        *err = "device index beyond the no-allocation table";
        return nullptr;
    }
    *dev_out = dev;
    std::lock_guard<std::mutex> lk(g_na_mutex);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        g_na_ptr[dev] = (char *)p;
    }
    return g_na_ptr[dev];
}

// clear n ints (a kernel, not a memset: safe under graph capture)
__global__ void k_fill_int(int *a, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// the argmax-y result (it_mark) of the shared cell kernels starts at -1
__global__ void k_na_start(int *misc)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        misc[0] = -1;  // mark
        misc[1] = 0;   // error bits
    }
}

// error bits of this call to the caller's device int
__global__ void k_na_status(const int *misc, int *d_err)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *d_err = misc[1];
}

}  // namespace

extern "C" int fipi_cells_build_noalloc(const fipi_params *params, const double *d_pos, int32_t n,
                                        int32_t *d_cell_start, int32_t *d_ids, int32_t *d_err,
                                        int32_t *ncell_out, void *stream)
{
    if (!params || n < 0 || (n > 0 && (!d_pos || !d_ids)) || !d_cell_start || !d_err)
        return set_err(nullptr, FIPI_ERR_ARG, "invalid argument");
    fipi_params q = *params;
    q.switch_particle = 1;
    q.n_particles = n;
    int rc = check_params(&q);
    if (rc != FIPI_OK) return rc;
    const PartParams base = part_params_of(q, n, 1);
    const int ncell = base.ncell_box;
    if (ncell_out) *ncell_out = ncell;
    if (n > NA_MAX_N)
        return set_err(nullptr, FIPI_ERR_ARG,
                       "n exceeds FIPI_NOALLOC_MAX_PARTICLES (" + std::to_string(NA_MAX_N) + ")");
    if (ncell > NA_MAX_CELLS)
        return set_err(nullptr, FIPI_ERR_ARG,
                       "the contact grid exceeds FIPI_NOALLOC_MAX_CELLS (" + std::to_string(NA_MAX_CELLS) + ")");
    int dev = 0;
    std::string msg;
    char *S = na_scratch(&dev, &msg);
    if (!S) return set_err(nullptr, FIPI_ERR_HIP, msg);
    hipStream_t st = (hipStream_t)stream;
    double *px = (double *)(S + NA_OFF_P);
    double *py = (double *)(S + NA_OFF_P + na_up(sizeof(double) * NA_MAX_N));
    double *pz = (double *)(S + NA_OFF_P + 2 * na_up(sizeof(double) * NA_MAX_N));
    int *cell = (int *)(S + NA_OFF_CELL), *cnt = (int *)(S + NA_OFF_CNT), *cur = (int *)(S + NA_OFF_CUR);
    int *bsum = (int *)(S + NA_OFF_BSUM), *yidx = (int *)(S + NA_OFF_YIDX), *misc = (int *)(S + NA_OFF_MISC);
    double *ymax = (double *)(S + NA_OFF_YMAX);
    // counts and cursors start at zero whatever an earlier call left
    hipLaunchKernelGGL(k_fill_int, grid1(ncell + 1, 256), dim3(256), 0, st, cnt, ncell + 1, 0);
    hipLaunchKernelGGL(k_fill_int, grid1(ncell + 1, 256), dim3(256), 0, st, cur, ncell + 1, 0);
    hipLaunchKernelGGL(k_na_start, dim3(1), dim3(64), 0, st, misc);
    if (n > 0) hipLaunchKernelGGL(k_aos_to_soa, grid1(n, 256), dim3(256), 0, st, d_pos, px, py, pz, n);
    PartState ps{};
    ps.p[0] = px;
    ps.p[1] = py;
    ps.p[2] = pz;
    ps.cell = cell;
    ps.err = misc + 1;
    ps.mark = misc;
    // the kernels of fipi_cells_build: histogram, scan, scatter, per-cell order
    hipLaunchKernelGGL(k_cell_count, grid1(n, 256), dim3(256), 0, st, base, (const double *)px, (const double *)py,
                       (const double *)pz, (const int *)nullptr, cell, cnt, ps.err, ymax, yidx);
    const int nb = ncell > 0 ? (ncell + SCAN_ITEMS - 1) / SCAN_ITEMS : 1;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, st, (const int *)cnt, d_cell_start, ncell, bsum);
    hipLaunchKernelGGL(k_scan_finish, dim3(nb), dim3(256), 0, st, d_cell_start, ncell, (const int *)bsum, nb,
                       (const double *)ymax, (const int *)yidx, (int)grid1(n, 256).x, misc);
    hipLaunchKernelGGL(k_cell_scatter, grid1(n, 256), dim3(256), 0, st, n, (const int *)cell,
                       (const int *)d_cell_start, cur, d_ids);
    hipLaunchKernelGGL(k_cell_sort, grid1(ncell, 256), dim3(256), 0, st, ncell, (const int *)d_cell_start, d_ids, ps,
                       cnt, cur);
    hipLaunchKernelGGL(k_na_status, dim3(1), dim3(64), 0, st, (const int *)misc, d_err);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(nullptr, FIPI_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
    return FIPI_OK;
}
