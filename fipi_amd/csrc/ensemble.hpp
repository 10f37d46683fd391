// Batched ensemble of independent periodic boxes (fipi_ensemble_* in
// include/fipi.h).  Included at the end of fipi.hip: it drives the same
// contexts, kernels and launch sequences as the single-box path.
//
// Boxes of equal edge length form one slab (a fipi_ctx with nbox > 1), so
// every pass of the step is one launch over all boxes of the slab
// (blockIdx.z = box).  Slabs run on their own streams.  After the first
// step (first-order, t = 0) a whole ensemble step is captured as one HIP
// graph per parity of the SBDF2 history ring (the only host state a step
// depends on once t > 0; the contact ramp is read from the device step
// counter), and each later step is one graph replay.

struct fipi_ensemble {
    fipi_params common{};
    std::string err;
    hipStream_t stream = nullptr;
    struct Slab {
        int n = 0;
        fipi_ctx *ctx = nullptr;
        std::vector<int> boxes;    // ensemble box of each slot
        std::vector<int> pfirst;   // particles of slot s: [pfirst[s], pfirst[s + 1])
        std::vector<double> phi, pos, rad, cth;
        std::vector<int> pbox;
        hipEvent_t done = nullptr;
    };
    std::vector<Slab> slabs;
    std::vector<std::pair<int, int>> where;  // box -> (slab, slot)
    std::vector<char> given;                 // initial state of the box set
    bool has_rad = false, has_cth = false;
    bool committed = false;
    hipEvent_t fork = nullptr;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int t = 0;
};

namespace {

int ens_err(fipi_ensemble *e, int code, const std::string &msg)
{
    if (e) e->err = msg;
    g_last_error = msg;
    return code;
}

#define ENSCHK(e, call)                                                                             \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return ens_err(e, FIPI_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

// a slab call failed: take over its message
int ens_from_ctx(fipi_ensemble *e, const fipi_ctx *ctx, int rc)
{
    return ens_err(e, rc, ctx ? ctx->err : g_last_error);
}

void ens_drop_graphs(fipi_ensemble *e)
{
    for (int k = 0; k < 2; k++) {
        if (e->exec[k]) (void)hipGraphExecDestroy(e->exec[k]);
        if (e->graph[k]) (void)hipGraphDestroy(e->graph[k]);
        e->exec[k] = nullptr;
        e->graph[k] = nullptr;
    }
}

void ens_step_slab(fipi_ctx *ctx)
{
    if (ctx->fp64) Impl<double>::step(ctx);
    else Impl<float>::step(ctx);
}

// Load every slab's initial state (t = 0).
int ens_commit(fipi_ensemble *e)
{
    for (size_t b = 0; b < e->given.size(); b++)
        if (!e->given[b]) return ens_err(e, FIPI_ERR_STATE, "box " + std::to_string(b) + " has no initial state");
    ENSCHK(e, hipStreamSynchronize(e->stream));
    ens_drop_graphs(e);
    for (auto &s : e->slabs) {
        const int np = s.pfirst.back();
        int rc = set_state_slab(s.ctx, s.phi.data(), np > 0 ? s.pos.data() : nullptr,
                                e->has_rad && np > 0 ? s.rad.data() : nullptr,
                                e->has_cth && np > 0 ? s.cth.data() : nullptr, np > 0 ? s.pbox.data() : nullptr);
        if (rc) return ens_from_ctx(e, s.ctx, rc);
    }
    e->t = 0;
    e->committed = true;
    return FIPI_OK;
}

// the ensemble stream waits for the slab streams
int ens_join(fipi_ensemble *e)
{
    for (auto &s : e->slabs) {
        ENSCHK(e, hipEventRecord(s.done, s.ctx->stream));
        ENSCHK(e, hipStreamWaitEvent(e->stream, s.done, 0));
    }
    return FIPI_OK;
}

// Capture one ensemble step per parity of the history ring.  The slabs'
// host state (t, cur) is set to what it will be at the replay and restored
// afterwards: capture records launches without running them.
int ens_capture(fipi_ensemble *e)
{
    std::vector<int> t0, c0;
    std::vector<bool> r0;
    for (auto &s : e->slabs) {
        t0.push_back(s.ctx->t);
        c0.push_back(s.ctx->cur);
        r0.push_back(s.ctx->conv_ready);
    }
    for (int k = 0; k < 2; k++) {
        for (size_t i = 0; i < e->slabs.size(); i++) {
            e->slabs[i].ctx->t = t0[i] + k;
            e->slabs[i].ctx->cur = c0[i] ^ k;
            e->slabs[i].ctx->conv_ready = r0[i];
        }
        ENSCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        hipError_t ce = hipEventRecord(e->fork, e->stream);
        for (auto &s : e->slabs) {
            if (ce == hipSuccess) ce = hipStreamWaitEvent(s.ctx->stream, e->fork, 0);
            if (ce == hipSuccess) ens_step_slab(s.ctx);
            if (ce == hipSuccess) ce = hipEventRecord(s.done, s.ctx->stream);
            if (ce == hipSuccess) ce = hipStreamWaitEvent(e->stream, s.done, 0);
        }
        hipGraph_t g = nullptr;
        hipError_t ee = hipStreamEndCapture(e->stream, &g);
        if (ce != hipSuccess || ee != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return ens_err(e, FIPI_ERR_HIP,
                           std::string("graph capture: ") + hipGetErrorString(ce != hipSuccess ? ce : ee));
        }
        e->graph[k] = g;
        ENSCHK(e, hipGraphInstantiate(&e->exec[k], g, nullptr, nullptr, 0));
    }
    for (size_t i = 0; i < e->slabs.size(); i++) {
        e->slabs[i].ctx->t = t0[i];
        e->slabs[i].ctx->cur = c0[i];
        e->slabs[i].ctx->conv_ready = r0[i];
    }
    return FIPI_OK;
}

}  // namespace

extern "C" {

int fipi_ensemble_create(const fipi_params *common, int32_t n_boxes, const fipi_box_desc *boxes,
                         fipi_ensemble **out)
{
    if (!out) return ens_err(nullptr, FIPI_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!common || n_boxes <= 0 || !boxes) return ens_err(nullptr, FIPI_ERR_ARG, "common, n_boxes > 0 and boxes required");
    if (common->nx <= 0 || !(common->lx > 0)) return ens_err(nullptr, FIPI_ERR_ARG, "common lx/nx gives the spacing");
    const double dx = common->lx / common->nx;
    auto *e = new fipi_ensemble();
    e->common = *common;
    e->where.resize(n_boxes);
    e->given.assign(n_boxes, 0);
    for (int b = 0; b < n_boxes; b++) {
        const int n = boxes[b].n;
        if (!supported_n(n) || boxes[b].n_particles < 0) {
            delete e;
            return ens_err(nullptr, FIPI_ERR_UNSUPPORTED, "box " + std::to_string(b) + ": unsupported edge or count");
        }
        size_t k = 0;
        while (k < e->slabs.size() && e->slabs[k].n != n) k++;
        if (k == e->slabs.size()) {
            e->slabs.emplace_back();
            e->slabs.back().n = n;
            e->slabs.back().pfirst.push_back(0);
        }
        auto &s = e->slabs[k];
        e->where[b] = {(int)k, (int)s.boxes.size()};
        s.boxes.push_back(b);
        s.pfirst.push_back(s.pfirst.back() + boxes[b].n_particles);
    }
    hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
    if (he != hipSuccess) {
        fipi_ensemble_destroy(e);
        return ens_err(nullptr, FIPI_ERR_HIP, std::string("HIP init: ") + hipGetErrorString(he));
    }
    for (auto &s : e->slabs) {
        fipi_params p = *common;
        p.nx = p.ny = p.nz = s.n;
        p.lx = p.ly = p.lz = s.n * dx;
        const int np = s.pfirst.back();
        p.n_particles = np;
        int rc = create_ctx(&p, (int)s.boxes.size(), np, &s.ctx);
        if (rc == FIPI_OK && hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
            rc = set_err(nullptr, FIPI_ERR_HIP, "hipEventCreate failed");
        if (rc != FIPI_OK) {
            const std::string msg = "edge " + std::to_string(s.n) + ": " + g_last_error;
            fipi_ensemble_destroy(e);
            return ens_err(nullptr, rc, msg);
        }
        const size_t G = (size_t)s.n * s.n * s.n;
        s.phi.assign(G * s.boxes.size(), 0.0);
        s.pos.assign(3 * (size_t)np, 0.0);
        s.rad.assign(np, common->r_p);
        s.cth.assign(np, 0.0);
        s.pbox.resize(np);
        for (size_t k = 0; k < s.boxes.size(); k++)
            for (int i = s.pfirst[k]; i < s.pfirst[k + 1]; i++) s.pbox[i] = (int)k;
    }
    *out = e;
    return FIPI_OK;
}

int fipi_ensemble_destroy(fipi_ensemble *e)
{
    if (!e) return FIPI_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    ens_drop_graphs(e);
    for (auto &s : e->slabs) {
        if (s.done) (void)hipEventDestroy(s.done);
        fipi_destroy(s.ctx);
    }
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return FIPI_OK;
}

const char *fipi_ensemble_last_error(const fipi_ensemble *e) { return e ? e->err.c_str() : g_last_error.c_str(); }

void *fipi_ensemble_stream(fipi_ensemble *e) { return e ? (void *)e->stream : nullptr; }

int fipi_ensemble_set_box(fipi_ensemble *e, int32_t box, const double *phi, const double *pos, const double *radius,
                          const double *cos_theta)
{
    if (!e || !phi || box < 0 || box >= (int)e->where.size()) return ens_err(e, FIPI_ERR_ARG, "invalid argument");
    auto &s = e->slabs[e->where[box].first];
    const int k = e->where[box].second;
    const int p0 = s.pfirst[k], np = s.pfirst[k + 1] - p0;
    if (np > 0 && !pos) return ens_err(e, FIPI_ERR_ARG, "pos is required for a box with particles");
    const size_t G = (size_t)s.n * s.n * s.n;
    std::copy(phi, phi + G, s.phi.begin() + (size_t)k * G);
    if (np > 0) std::copy(pos, pos + 3 * (size_t)np, s.pos.begin() + 3 * (size_t)p0);
    for (int i = 0; i < np; i++) {
        s.rad[p0 + i] = radius ? radius[i] : e->common.r_p;
        s.cth[p0 + i] = cos_theta ? cos_theta[i] : 0.0;
    }
    if (radius && np > 0) e->has_rad = true;
    if (cos_theta && np > 0) e->has_cth = true;
    e->given[box] = 1;
    e->committed = false;
    return FIPI_OK;
}

int fipi_ensemble_tune(fipi_ensemble *e, int32_t steps)
{
    if (!e || steps < 1) return ens_err(e, FIPI_ERR_ARG, "invalid argument");
    int rc = ens_commit(e);
    if (rc) return rc;
    struct Events {
        hipEvent_t a = nullptr, b = nullptr;
        ~Events()
        {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    ENSCHK(e, hipEventCreate(&ev.a));
    ENSCHK(e, hipEventCreate(&ev.b));
    // per slab, time every combination; the first error ends the trials
    auto trials = [&]() -> int {
        for (auto &s : e->slabs) {
            fipi_ctx *ctx = s.ctx;
            fipi_ctx::Tune best = ctx->tune;
            float best_ms = 1e30f;
            int ydef, yalt;
            y_plans_of(ctx, &ydef, &yalt);
            std::vector<int> plans = {0};  // 0: the default plan of the size
            if (yalt != 0) plans.push_back(yalt);
            for (int ye : plans)
                for (int y : {64, 128})
                    for (int h : {64, 128})
                        for (int st : {64, 128}) {
                            ctx->tune = fipi_ctx::Tune{y, h, st, ye};
                            ens_step_slab(ctx);  // warm-up (first use of these kernels)
                            ENSCHK(e, hipEventRecord(ev.a, ctx->stream));
                            for (int k = 0; k < steps; k++) ens_step_slab(ctx);
                            ENSCHK(e, hipEventRecord(ev.b, ctx->stream));
                            ENSCHK(e, hipEventSynchronize(ev.b));
                            float ms = 0.f;
                            ENSCHK(e, hipEventElapsedTime(&ms, ev.a, ev.b));
                            if (ms < best_ms) {
                                best_ms = ms;
                                best = ctx->tune;
                            }
                        }
            ctx->tune = best;
            const int frc = check_flags(ctx);
            if (frc) return ens_from_ctx(e, ctx, frc);
        }
        return FIPI_OK;
    };
    rc = trials();
    const int crc = ens_commit(e);  // back to the initial state, also after a failed trial
    return rc ? rc : crc;
}

int fipi_ensemble_get_tuning(const fipi_ensemble *e, int32_t n, fipi_tuning_t *out)
{
    if (!e || !out) return ens_err(nullptr, FIPI_ERR_ARG, "ens and out are required");
    *out = fipi_tuning_t{0, 0, 0, 0};
    for (const auto &s : e->slabs)
        if (s.n == n) tuning_out(s.ctx, out);
    return FIPI_OK;
}

int fipi_ensemble_set_tuning(fipi_ensemble *e, int32_t n, const fipi_tuning_t *in)
{
    if (!e || !in) return ens_err(e, FIPI_ERR_ARG, "ens and in are required");
    for (auto &s : e->slabs)
        if (s.n == n) {
            ENSCHK(e, hipStreamSynchronize(e->stream));
            const int rc = tuning_in(s.ctx, in);
            if (rc) return ens_from_ctx(e, s.ctx, rc);
            ens_drop_graphs(e);  // the captured launches used the old choice
            return FIPI_OK;
        }
    return ens_err(e, FIPI_ERR_ARG, "no box of that edge length");
}

int fipi_ensemble_step_async(fipi_ensemble *e, int32_t nsteps)
{
    if (!e || nsteps < 0) return ens_err(e, FIPI_ERR_ARG, "invalid argument");
    int rc;
    if (!e->committed && (rc = ens_commit(e))) return rc;
    for (int k = 0; k < nsteps; k++) {
        if (e->t == 0) {
            // first-order start (main.cpp:95-114): launched directly
            for (auto &s : e->slabs) ens_step_slab(s.ctx);
            if ((rc = ens_join(e))) return rc;
        } else {
            if (!e->exec[0] && (rc = ens_capture(e))) return rc;
            ENSCHK(e, hipGraphLaunch(e->exec[(e->t - 1) & 1], e->stream));
            for (auto &s : e->slabs) {
                s.ctx->cur ^= 1;
                s.ctx->t++;
            }
        }
        e->t++;
    }
    ENSCHK(e, hipGetLastError());
    return FIPI_OK;
}

int fipi_ensemble_sync(fipi_ensemble *e)
{
    if (!e) return ens_err(e, FIPI_ERR_ARG, "ensemble is NULL");
    ENSCHK(e, hipStreamSynchronize(e->stream));
    for (auto &s : e->slabs) {
        int rc = check_flags(s.ctx);
        if (rc) return ens_from_ctx(e, s.ctx, rc);
    }
    return FIPI_OK;
}

int fipi_ensemble_step(fipi_ensemble *e, int32_t nsteps)
{
    int rc = fipi_ensemble_step_async(e, nsteps);
    if (rc) return rc;
    return fipi_ensemble_sync(e);
}

int fipi_ensemble_get_phi(fipi_ensemble *e, int32_t box, double *phi)
{
    if (!e || !phi || box < 0 || box >= (int)e->where.size()) return ens_err(e, FIPI_ERR_ARG, "invalid argument");
    if (!e->committed) return ens_err(e, FIPI_ERR_STATE, "no state yet");
    ENSCHK(e, hipStreamSynchronize(e->stream));
    auto &s = e->slabs[e->where[box].first];
    const size_t G = s.ctx->G;
    const void *src = (const char *)s.ctx->c + (size_t)e->where[box].second * G * s.ctx->esz;
    int rc = copy_out_real(s.ctx, phi, src, G);
    return rc ? ens_from_ctx(e, s.ctx, rc) : FIPI_OK;
}

int fipi_ensemble_get_particles(fipi_ensemble *e, int32_t box, double *pos)
{
    if (!e || !pos || box < 0 || box >= (int)e->where.size()) return ens_err(e, FIPI_ERR_ARG, "invalid argument");
    if (!e->committed) return ens_err(e, FIPI_ERR_STATE, "no state yet");
    ENSCHK(e, hipStreamSynchronize(e->stream));
    auto &s = e->slabs[e->where[box].first];
    const int k = e->where[box].second, p0 = s.pfirst[k], np = s.pfirst[k + 1] - p0;
    if (np == 0) return FIPI_OK;
    fipi_ctx *ctx = s.ctx;
    double *tmp;
    ENSCHK(e, hipMalloc((void **)&tmp, sizeof(double) * 3 * np));
    hipLaunchKernelGGL(k_soa_to_aos, grid1(np, 256), dim3(256), 0, ctx->stream, ctx->ps.p[0] + p0, ctx->ps.p[1] + p0,
                       ctx->ps.p[2] + p0, tmp, np);
    hipError_t he = hipMemcpyAsync(pos, tmp, sizeof(double) * 3 * np, hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    if (he != hipSuccess) return ens_err(e, FIPI_ERR_HIP, hipGetErrorString(he));
    return FIPI_OK;
}

int fipi_ensemble_diagnostics(fipi_ensemble *e, int32_t box, fipi_diag *out)
{
    if (!e || !out || box < 0 || box >= (int)e->where.size()) return ens_err(e, FIPI_ERR_ARG, "invalid argument");
    if (!e->committed) return ens_err(e, FIPI_ERR_STATE, "no state yet");
    ENSCHK(e, hipStreamSynchronize(e->stream));
    auto &s = e->slabs[e->where[box].first];
    const int k = e->where[box].second;
    int rc = diag_box(s.ctx, k, s.pfirst[k], s.pfirst[k + 1], out);
    if (rc) return ens_from_ctx(e, s.ctx, rc);
    out->t = e->t;
    return FIPI_OK;
}

}  // extern "C"
