// Hand-written adjoint of the CH/Stokes step for calibration
// (fipi_adjoint_* in include/fipi.h).  Included at the end of fipi.hip.
//
// Model (particles held fixed, rho1 = rho2): the step of src/main.cpp:75-222
// with a wetting source in the chemical potential,
//   f(c) = (lambda/eps^2)(c^3 - c) + W,  W = -a_w sum_p cos(theta_p) S_p,
// S_p the trilinear spreading kernel of particle p divided by dx^3 (the
// transpose of field::interpolate, field.cpp:223-248; zero at theta = 90
// deg, where the step is the reference step).  In K-space, with
// L = k^2 (-Laplacian), D = 3 + 2 dt M lambda k^4, D0 = 1 + dt M lambda k^4,
//   G(c)   = -dt M L f(c) - dt N(c),   N(c) = U(c) . grad c,
//   U(c)   = S[mu(c) grad c + F_c],     mu = lambda L c + f(c),
// S the Stokes operator (projection / (eta k^2)) and F_c the constant
// spread contact force of the fixed particles:
//   c^1     = D0^-1 [c^0 + G0(c^0)]              (G0 = G without N: u^0 = 0)
//   c^{n+1} = D^-1 [4c^n + 4G(c^n) - c^{n-1} - 2G(c^{n-1})],  n >= 1.
// Loss J = sum_x (c^T - c*)^2.  With abar^n = dJ/dc^n, b^m = D^-1 abar^m
// (b0 = D0^-1 abar^1) and r^n = 4b^{n+1} - 2b^{n+2} (r^0 = b0 - 2b^2):
//   abar^n = 4b^{n+1} - b^{n+2} + J_G(c^n)^T r^n        (n >= 1)
//   J_G^T r = -dt M f'(c) L r - dt J_N^T r,
//   J_N^T r = -div(r U) + lambda L q + f'(c) q - div(mu z),
//   z = S(r grad c), q = z . grad c.
// Gradients (sums over n of the direct dependence of step n+1):
//   dJ/dM      += <r^n, -dt L f>  -  <b^{n+1}, dD/dM c^{n+1}>
//   dJ/dlambda += <r^n, -dt M L f0 - dt N_lambda> - <b^{n+1}, dD/dlambda c^{n+1}>
//                 with f0 = (c^3 - c)/eps^2, <r, N_lambda> = <q, L c + f0>
//   dJ/dtheta_p += a_w sin(theta_p) interp(psi^n, x_p) / dx^3,
//                 psi^n = -dt M L r^n - dt q^n
// (the adjoint of spreading is trilinear interpolation).  Gradient,
// divergence and L are the forward's spectral operators (Nyquist planes of
// a component zeroed as in field.cpp:310-318), so their transposes are
// exact.  The forward is the fipi step itself; checkpoints of its state
// every K steps are restored and each segment recomputed for the backward
// sweep.

namespace {

// ---- spectral operators on a K-space half spectrum H[z][y][kx] ----------
// one thread per (kx < nxh, y, z); s folds the inverse transform's 1/G
template <typename T>
__global__ void k_spec_scale(cplx<T> *a, Dims d, const double *kxv, const double *kyv, const double *kzv, int op,
                             double c1, double s)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)d.nxh * d.ny * d.nz;
    if (i >= n) return;
    const int kx = (int)(i % d.nxh), y = (int)((i / d.nxh) % d.ny), z = (int)(i / ((size_t)d.nxh * d.ny));
    const double k2 = kxv[kx] * kxv[kx] + kyv[y] * kyv[y] + kzv[z] * kzv[z];
    double m;
    if (op == 0) m = k2 * s;                          // L
    else if (op == 1) m = s / (3.0 + c1 * k2 * k2);   // D^-1, c1 = 2 dt M lambda
    else if (op == 2) m = s / (1.0 + c1 * k2 * k2);   // D0^-1, c1 = dt M lambda
    else m = s;
    const size_t j = hidx(d, kx, y, z);
    a[j] = cplx<T>{(T)((double)a[j].x * m), (T)((double)a[j].y * m)};
}

// out = i k_axis chi a * s (gradient component; zero on the axis' Nyquist plane)
template <typename T>
__global__ void k_spec_grad(const cplx<T> *a, cplx<T> *out, Dims d, const double *kv, int axis, double s)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)d.nxh * d.ny * d.nz;
    if (i >= n) return;
    const int kx = (int)(i % d.nxh), y = (int)((i / d.nxh) % d.ny), z = (int)(i / ((size_t)d.nxh * d.ny));
    const int idx = axis == 0 ? kx : (axis == 1 ? y : z);
    const int nn = axis == 0 ? d.nx : (axis == 1 ? d.ny : d.nz);
    const double k = idx == nn / 2 ? 0.0 : kv[idx] * s;
    const size_t j = hidx(d, kx, y, z);
    out[j] = cplx<T>{(T)(-k * (double)a[j].y), (T)(k * (double)a[j].x)};
}

// out = sum_d i k_d chi_d a_d * s (divergence)
template <typename T>
__global__ void k_spec_div3(const cplx<T> *a0, const cplx<T> *a1, const cplx<T> *a2, cplx<T> *out, Dims d,
                            const double *kxv, const double *kyv, const double *kzv, double s)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)d.nxh * d.ny * d.nz;
    if (i >= n) return;
    const int kx = (int)(i % d.nxh), y = (int)((i / d.nxh) % d.ny), z = (int)(i / ((size_t)d.nxh * d.ny));
    const double k0 = kx == d.nx / 2 ? 0.0 : kxv[kx], k1 = y == d.ny / 2 ? 0.0 : kyv[y], k2 = z == d.nz / 2 ? 0.0 : kzv[z];
    const size_t j = hidx(d, kx, y, z);
    const double re = -(k0 * (double)a0[j].y + k1 * (double)a1[j].y + k2 * (double)a2[j].y);
    const double im = k0 * (double)a0[j].x + k1 * (double)a1[j].x + k2 * (double)a2[j].x;
    out[j] = cplx<T>{(T)(re * s), (T)(im * s)};
}

// Stokes operator in place: (F - k (k.F)/k^2) / (eta k^2) * s, zero at k = 0,
// with the forward's Nyquist-class split of k.F (k_hub_stokes)
template <typename T>
__global__ void k_spec_stokes3(cplx<T> *a0, cplx<T> *a1, cplx<T> *a2, Dims d, const double *kxv, const double *kyv,
                               const double *kzv, double inv_mu_s)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)d.nxh * d.ny * d.nz;
    if (i >= n) return;
    const int kx = (int)(i % d.nxh), y = (int)((i / d.nxh) % d.ny), z = (int)(i / ((size_t)d.nxh * d.ny));
    const double kv[3] = {kxv[kx], kyv[y], kzv[z]};
    const bool nq[3] = {kx == d.nx / 2, y == d.ny / 2, z == d.nz / 2};
    const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
    const size_t j = hidx(d, kx, y, z);
    cplx<T> *a[3] = {a0, a1, a2};
    if (k2 == 0.0) {
        for (int e = 0; e < 3; e++) a[e][j] = cplx<T>{T(0), T(0)};
        return;
    }
    double F[3][2], sar = 0, sai = 0, sbr = 0, sbi = 0;
    for (int e = 0; e < 3; e++) {
        F[e][0] = a[e][j].x;
        F[e][1] = a[e][j].y;
        if (nq[e]) { sbr += kv[e] * F[e][0]; sbi += kv[e] * F[e][1]; }
        else { sar += kv[e] * F[e][0]; sai += kv[e] * F[e][1]; }
    }
    const double ik2 = 1.0 / k2, cu = ik2 * inv_mu_s;
    for (int e = 0; e < 3; e++) {
        const double sr = nq[e] ? sbr : sar, si = nq[e] ? sbi : sai;
        a[e][j] = cplx<T>{(T)((F[e][0] - kv[e] * sr * ik2) * cu), (T)((F[e][1] - kv[e] * si * ik2) * cu)};
    }
}

// Partial sums of w(kx) k^4 Re(a conj b) (w = 1 on the kx = 0 and nx/2
// columns, 2 elsewhere): <L A, L B> over the full real grid times G
template <typename T>
__global__ __launch_bounds__(256) void k_spec_dot_k4(const cplx<T> *a, const cplx<T> *b, Dims d, const double *kxv,
                                                     const double *kyv, const double *kzv, double *part)
{
    __shared__ double s[256];
    const size_t n = (size_t)d.nxh * d.ny * d.nz;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int kx = (int)(i % d.nxh), y = (int)((i / d.nxh) % d.ny), z = (int)(i / ((size_t)d.nxh * d.ny));
        const double k2 = kxv[kx] * kxv[kx] + kyv[y] * kyv[y] + kzv[z] * kzv[z];
        const double w = (kx == 0 || 2 * kx == d.nx) ? 1.0 : 2.0;
        const size_t j = hidx(d, kx, y, z);
        acc += w * k2 * k2 * ((double)a[j].x * (double)b[j].x + (double)a[j].y * (double)b[j].y);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Partial sums of a.b (or of (a - b)^2 when sq) over n reals
template <typename T>
__global__ __launch_bounds__(256) void k_dot_partial(const T *a, const T *b, size_t n, int sq, double *part)
{
    __shared__ double s[256];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double x = a[i], y = b[i];
        acc += sq ? (x - y) * (x - y) : x * y;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// ---- pointwise kernels over n reals -------------------------------------
// f = sf (c^3 - c) + W, f' = sf (3c^2 - 1), f0 = (c^3 - c) / eps^2
template <typename T>
__global__ void k_ad_f(const T *c, const T *W, T *f, T *fp, T *f0, size_t n, T sf, double inv_eps2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T v = c[i];
    T r = (v * v * v - v) * sf;
    if (W) r += W[i];
    f[i] = r;
    fp[i] = (T((double)3 * v * v) - T(1)) * sf;
    f0[i] = (T)(((double)v * v * v - v) * inv_eps2);
}

// out = al a + be b + ga c (null operands count zero)
template <typename T>
__global__ void k_axpbypc(T *out, const T *a, double al, const T *b, double be, const T *c, double ga, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r = 0.0;
    if (a) r += al * (double)a[i];
    if (b) r += be * (double)b[i];
    if (c) r += ga * (double)c[i];
    out[i] = (T)r;
}

// o_d = a g_d (+ b_d when b0 is set: o_d = a g_d + m h_d)
template <typename T>
__global__ void k_mul3(const T *a, const T *g0, const T *g1, const T *g2, const T *m, const T *h0, const T *h1,
                       const T *h2, const T *e0, const T *e1, const T *e2, T *o0, T *o1, T *o2, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double av = a[i];
    double r0 = av * (double)g0[i], r1 = av * (double)g1[i], r2 = av * (double)g2[i];
    if (m) {
        const double mv = m[i];
        r0 += mv * (double)h0[i];
        r1 += mv * (double)h1[i];
        r2 += mv * (double)h2[i];
    }
    if (e0) {
        r0 += (double)e0[i];
        r1 += (double)e1[i];
        r2 += (double)e2[i];
    }
    o0[i] = (T)r0;
    o1[i] = (T)r1;
    o2[i] = (T)r2;
}

// q = z . g
template <typename T>
__global__ void k_dot3(const T *z0, const T *z1, const T *z2, const T *g0, const T *g1, const T *g2, T *q, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    q[i] = (T)((double)z0[i] * g0[i] + (double)z1[i] * g1[i] + (double)z2[i] * g2[i]);
}

// J_G^T r = -dtM f' Lr - dt (-divV + lambda Lq + f' q);  psi = -dtM Lr - dt q
// (conv off: the q, Lq, divV terms are absent)
template <typename T>
__global__ void k_ad_jg(const T *fp, const T *Lr, const T *divV, const T *Lq, const T *q, T *jg, T *psi, size_t n,
                        double dtM, double dt, double lam, int conv)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double f1 = fp[i], lr = Lr[i];
    double j = -dtM * f1 * lr, p = -dtM * lr;
    if (conv) {
        const double qq = q[i];
        j -= dt * (-(double)divV[i] + lam * (double)Lq[i] + f1 * qq);
        p -= dt * qq;
    }
    jg[i] = (T)j;
    psi[i] = (T)p;
}

// g_theta[p] += a sin(theta_p) v[p]
__global__ void k_theta_acc(double *g, const double *theta, const double *v, int n, double a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) g[i] += a * sin(theta[i]) * v[i];
}

// w[p] = -a cos(theta_p)
__global__ void k_wet_weight(const double *theta, double *w, int n, double a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = -a * cos(theta[i]);
}

}  // namespace

struct fipi_adjoint {
    std::string err;
    fipi_ctx *ctx = nullptr;
    int steps = 0, every = 1;
    double a_w = 0.0;
    int n = 0;
    std::vector<double> phi0, pos0;
    bool has_target = false;
    // device arrays (context precision unless noted)
    std::vector<void *> mem;
    void *target = nullptr, *W = nullptr, *spread3 = nullptr;
    double *dpos = nullptr;  // particle positions [n][3]
    double *theta = nullptr, *wgt = nullptr, *zero = nullptr, *gth = nullptr, *pv = nullptr;
    double *part = nullptr, *acc = nullptr;
    // checkpoints: per checkpoint one block of state buffers
    std::vector<std::vector<void *>> ckpt;
    std::vector<int> ck_t, ck_cur;
    std::vector<bool> ck_ready;
    std::vector<void *> seg;  // real c^n of the current segment
    // work arrays
    enum {
        C_, F_, FP_, F0_, MU_, LC_, G0_, G1_, G2_, U0_, U1_, U2_, R_, LR_, Z0_, Z1_, Z2_, Q_, LQ_, V0_, V1_, V2_,
        DIV_, JG_, PSI_, B1_, B2_, B0_, AB_, FC0_, FC1_, FC2_, NREAL
    };
    void *rw[NREAL] = {};
    void *hw[4] = {};  // half spectra
};

namespace {

int adj_err(fipi_adjoint *a, int code, const std::string &msg)
{
    if (a) a->err = msg;
    g_last_error = msg;
    return code;
}

#define ADJCHK(a, call)                                                                             \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return adj_err(a, FIPI_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

int adj_alloc(fipi_adjoint *a, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return adj_err(a, FIPI_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    a->mem.push_back(*p);
    return FIPI_OK;
}

// state buffers of the context a checkpoint holds, with their sizes
std::vector<std::pair<void *, size_t>> state_buffers(fipi_ctx *ctx)
{
    const size_t R = ctx->G * ctx->esz, H = ctx->GH * 2 * ctx->esz;
    return {{ctx->c, R},          {ctx->g[0], R},   {ctx->g[1], R},   {ctx->g[2], R},
            {ctx->u[0], R},       {ctx->u[1], R},   {ctx->u[2], R},   {ctx->ck[0], H},
            {ctx->ck[1], H},      {ctx->fk, H},     {ctx->hist, H},   {ctx->h[0], H},
            {ctx->k0, sizeof(double) * K0_STRIDE}, {ctx->conv0, sizeof(double) * 2},
            {ctx->conv_part, sizeof(double) * (size_t)(ctx->fp64 ? xblocks<double>(ctx->d) : xblocks<float>(ctx->d))},
            {ctx->t_dev, sizeof(int)}};
}

template <typename T>
struct Adj {
    using C = cplx<T>;
    static T *Rp(fipi_adjoint *a, int k) { return (T *)a->rw[k]; }
    static C *Hp(fipi_adjoint *a, int k) { return (C *)a->hw[k]; }

    // real -> K-space (unnormalised)
    static void fwd(fipi_ctx *ctx, const T *in, C *out)
    {
        Impl<T>::template x_forward<1>(ctx, ProdCopy<T>{in}, out, nullptr);
        Impl<T>::template y_pass<-1, 1, false>(ctx, out, nullptr, nullptr, nullptr);
        Impl<T>::template z_pass<-1>(ctx, out);
    }
    // K-space -> real (unnormalised; the spectral operator folds 1/G); destroys h
    static void inv(fipi_ctx *ctx, C *h, T *out)
    {
        Impl<T>::template z_pass<+1>(ctx, h);
        Impl<T>::template y_pass<+1, 1, false>(ctx, h, nullptr, nullptr, nullptr);
        Impl<T>::template x_inverse<1>(ctx, h, nullptr, nullptr, out, nullptr, nullptr);
    }
    static dim3 hg(fipi_ctx *ctx) { return grid1((size_t)ctx->d.nxh * ctx->d.ny * ctx->d.nz, 256); }
    static dim3 rg(fipi_ctx *ctx) { return grid1(ctx->G, 256); }

    // out = IFFT(k^2 FFT(in))  (L)
    static void lap(fipi_adjoint *a, const T *in, T *out, double scale = 1.0)
    {
        fipi_ctx *ctx = a->ctx;
        C *h = Hp(a, 0);
        fwd(ctx, in, h);
        hipLaunchKernelGGL((k_spec_scale<T>), hg(ctx), dim3(256), 0, ctx->stream, h, ctx->d, ctx->kx, ctx->ky, ctx->kz,
                           0, 0.0, scale / (double)ctx->G);
        inv(ctx, h, out);
    }
    // o_d = S(i_d) (3 fields, real -> real)
    static void stokes(fipi_adjoint *a, const T *i0, const T *i1, const T *i2, T *o0, T *o1, T *o2)
    {
        fipi_ctx *ctx = a->ctx;
        C *h0 = Hp(a, 0), *h1 = Hp(a, 1), *h2 = Hp(a, 2);
        fwd(ctx, i0, h0);
        fwd(ctx, i1, h1);
        fwd(ctx, i2, h2);
        const double inv_mu = ctx->p.mu != 0.0 ? 1.0 / ctx->p.mu : 0.0;
        hipLaunchKernelGGL((k_spec_stokes3<T>), hg(ctx), dim3(256), 0, ctx->stream, h0, h1, h2, ctx->d, ctx->kx,
                           ctx->ky, ctx->kz, inv_mu / (double)ctx->G);
        inv(ctx, h0, o0);
        inv(ctx, h1, o1);
        inv(ctx, h2, o2);
    }
    // acc slot <- <x, y> (or sum (x - y)^2), fixed order
    static void dot(fipi_adjoint *a, const T *x, const T *y, int sq, double *slot)
    {
        fipi_ctx *ctx = a->ctx;
        hipLaunchKernelGGL((k_dot_partial<T>), dim3(DIAG_BLOCKS), dim3(256), 0, ctx->stream, x, y, ctx->G, sq, a->part);
        hipLaunchKernelGGL(k_sum_strided, dim3(1), dim3(256), 0, ctx->stream, (const double *)a->part, DIAG_BLOCKS, 1,
                           1, slot);
    }

    // dense spread of per-particle vectors (w0, w1, w2) into spread3
    static void spread_dense(fipi_adjoint *a, const double *w0, const double *w1, const double *w2)
    {
        fipi_ctx *ctx = a->ctx;
        (void)hipMemsetAsync(a->spread3, 0, 3 * ctx->G * ctx->esz, ctx->stream);
        PartParams pp = ctx->pp;
        const PartState &ps = ctx->ps;
        Impl<T>::spread_build(ctx, pp, ps.p[0], ps.p[1], ps.p[2], w0, w1, w2, nullptr);
        hipLaunchKernelGGL((k_spread_dense<T>), grid1(ctx->nlines, 256), dim3(256), 0, ctx->stream, ctx->nlines,
                           ctx->p.nx, ctx->G, ctx->line_start, (const int *)ctx->node_cnt,
                           (const SpreadNode<T> *)ctx->spr_tmp, (T *)a->spread3);
    }

    // W = -a_w sum_p cos(theta_p) S_p (dense spread of the weights)
    static void wetting(fipi_adjoint *a)
    {
        fipi_ctx *ctx = a->ctx;
        if (a->n == 0 || a->a_w == 0.0) {
            (void)hipMemsetAsync(a->W, 0, ctx->G * ctx->esz, ctx->stream);
            return;
        }
        hipLaunchKernelGGL(k_wet_weight, grid1(a->n, 256), dim3(256), 0, ctx->stream, (const double *)a->theta, a->wgt,
                           a->n, a->a_w);
        spread_dense(a, a->wgt, a->zero, a->zero);
        (void)hipMemcpyAsync(a->W, a->spread3, ctx->G * ctx->esz, hipMemcpyDeviceToDevice, ctx->stream);
    }

    // F_c: the spread contact forces of the fixed particles (constant)
    static void contact_field(fipi_adjoint *a)
    {
        fipi_ctx *ctx = a->ctx;
        const PartState &ps = ctx->ps;
        spread_dense(a, ps.fpp[0], ps.fpp[1], ps.fpp[2]);
        const size_t R = ctx->G * ctx->esz;
        for (int d = 0; d < 3; d++)
            (void)hipMemcpyAsync(a->rw[fipi_adjoint::FC0_ + d], (const char *)a->spread3 + d * R, R,
                                 hipMemcpyDeviceToDevice, ctx->stream);
    }

    // Forward quantities at c (real, in C_): f, f', f0, grad c, mu, L c and,
    // with conv, U = S(mu grad c + F_c).
    static void forward_at(fipi_adjoint *a, bool conv)
    {
        fipi_ctx *ctx = a->ctx;
        hipStream_t st = ctx->stream;
        const size_t G = ctx->G;
        const double lam = ctx->p.lambda, eps = ctx->p.epsilon;
        hipLaunchKernelGGL((k_ad_f<T>), rg(ctx), dim3(256), 0, st, (const T *)Rp(a, a->C_), (const T *)a->W,
                           Rp(a, a->F_), Rp(a, a->FP_), Rp(a, a->F0_), G, (T)(lam / (eps * eps)), 1.0 / (eps * eps));
        C *hc = Hp(a, 3);
        fwd(ctx, Rp(a, a->C_), hc);
        const double *kv[3] = {ctx->kx, ctx->ky, ctx->kz};
        const int gi[3] = {a->G0_, a->G1_, a->G2_};
        for (int d = 0; d < 3; d++) {
            hipLaunchKernelGGL((k_spec_grad<T>), hg(ctx), dim3(256), 0, st, (const C *)hc, Hp(a, 0), ctx->d, kv[d], d,
                               1.0 / (double)G);
            inv(ctx, Hp(a, 0), Rp(a, gi[d]));
        }
        hipLaunchKernelGGL((k_spec_scale<T>), hg(ctx), dim3(256), 0, st, hc, ctx->d, ctx->kx, ctx->ky, ctx->kz, 0, 0.0,
                           1.0 / (double)G);
        inv(ctx, hc, Rp(a, a->LC_));
        // mu = lambda L c + f
        hipLaunchKernelGGL((k_axpbypc<T>), rg(ctx), dim3(256), 0, st, Rp(a, a->MU_), (const T *)Rp(a, a->LC_), lam,
                           (const T *)Rp(a, a->F_), 1.0, (const T *)nullptr, 0.0, G);
        if (conv) {
            // F = mu grad c + F_c -> U
            hipLaunchKernelGGL((k_mul3<T>), rg(ctx), dim3(256), 0, st, (const T *)Rp(a, a->MU_),
                               (const T *)Rp(a, a->G0_), (const T *)Rp(a, a->G1_), (const T *)Rp(a, a->G2_),
                               (const T *)nullptr, (const T *)nullptr, (const T *)nullptr, (const T *)nullptr,
                               (const T *)Rp(a, a->FC0_), (const T *)Rp(a, a->FC1_), (const T *)Rp(a, a->FC2_),
                               Rp(a, a->V0_), Rp(a, a->V1_), Rp(a, a->V2_), G);
            stokes(a, Rp(a, a->V0_), Rp(a, a->V1_), Rp(a, a->V2_), Rp(a, a->U0_), Rp(a, a->U1_), Rp(a, a->U2_));
        }
    }

    // J_G(c)^T r (r in R_) -> JG_, psi -> PSI_, q -> Q_; gradient dots of
    // step n into acc[5n..5n+2]; theta gradients accumulated.
    static void adjoint_at(fipi_adjoint *a, bool conv, double *acc_n)
    {
        fipi_ctx *ctx = a->ctx;
        hipStream_t st = ctx->stream;
        const size_t G = ctx->G;
        const double dt = ctx->p.dt, M = ctx->p.mobility, lam = ctx->p.lambda;
        forward_at(a, conv);
        lap(a, Rp(a, a->R_), Rp(a, a->LR_));
        if (conv) {
            // z = S(r grad c), q = z . grad c, L q, div(r U + mu z)
            hipLaunchKernelGGL((k_mul3<T>), rg(ctx), dim3(256), 0, st, (const T *)Rp(a, a->R_),
                               (const T *)Rp(a, a->G0_), (const T *)Rp(a, a->G1_), (const T *)Rp(a, a->G2_),
                               (const T *)nullptr, (const T *)nullptr, (const T *)nullptr, (const T *)nullptr,
                               (const T *)nullptr, (const T *)nullptr, (const T *)nullptr, Rp(a, a->V0_),
                               Rp(a, a->V1_), Rp(a, a->V2_), G);
            stokes(a, Rp(a, a->V0_), Rp(a, a->V1_), Rp(a, a->V2_), Rp(a, a->Z0_), Rp(a, a->Z1_), Rp(a, a->Z2_));
            hipLaunchKernelGGL((k_dot3<T>), rg(ctx), dim3(256), 0, st, (const T *)Rp(a, a->Z0_),
                               (const T *)Rp(a, a->Z1_), (const T *)Rp(a, a->Z2_), (const T *)Rp(a, a->G0_),
                               (const T *)Rp(a, a->G1_), (const T *)Rp(a, a->G2_), Rp(a, a->Q_), G);
            lap(a, Rp(a, a->Q_), Rp(a, a->LQ_));
            hipLaunchKernelGGL((k_mul3<T>), rg(ctx), dim3(256), 0, st, (const T *)Rp(a, a->R_),
                               (const T *)Rp(a, a->U0_), (const T *)Rp(a, a->U1_), (const T *)Rp(a, a->U2_),
                               (const T *)Rp(a, a->MU_), (const T *)Rp(a, a->Z0_), (const T *)Rp(a, a->Z1_),
                               (const T *)Rp(a, a->Z2_), (const T *)nullptr, (const T *)nullptr, (const T *)nullptr,
                               Rp(a, a->V0_), Rp(a, a->V1_), Rp(a, a->V2_), G);
            C *h0 = Hp(a, 0), *h1 = Hp(a, 1), *h2 = Hp(a, 2);
            fwd(ctx, Rp(a, a->V0_), h0);
            fwd(ctx, Rp(a, a->V1_), h1);
            fwd(ctx, Rp(a, a->V2_), h2);
            hipLaunchKernelGGL((k_spec_div3<T>), hg(ctx), dim3(256), 0, st, (const C *)h0, (const C *)h1,
                               (const C *)h2, Hp(a, 3), ctx->d, ctx->kx, ctx->ky, ctx->kz, 1.0 / (double)G);
            inv(ctx, Hp(a, 3), Rp(a, a->DIV_));
        }
        hipLaunchKernelGGL((k_ad_jg<T>), rg(ctx), dim3(256), 0, st, (const T *)Rp(a, a->FP_), (const T *)Rp(a, a->LR_),
                           (const T *)Rp(a, a->DIV_), (const T *)Rp(a, a->LQ_), (const T *)Rp(a, a->Q_),
                           Rp(a, a->JG_), Rp(a, a->PSI_), G, dt * M, dt, lam, conv ? 1 : 0);
        // <Lr, f>, <Lr, f0>, <q, Lc + f0>
        dot(a, Rp(a, a->LR_), Rp(a, a->F_), 0, acc_n + 0);
        dot(a, Rp(a, a->LR_), Rp(a, a->F0_), 0, acc_n + 1);
        if (conv) {
            hipLaunchKernelGGL((k_axpbypc<T>), rg(ctx), dim3(256), 0, st, Rp(a, a->V0_), (const T *)Rp(a, a->LC_), 1.0,
                               (const T *)Rp(a, a->F0_), 1.0, (const T *)nullptr, 0.0, G);
            dot(a, Rp(a, a->Q_), Rp(a, a->V0_), 0, acc_n + 2);
        } else {
            (void)hipMemsetAsync(acc_n + 2, 0, sizeof(double), st);
        }
        // theta: a_w sin(theta) interp(psi) / dx^3
        if (a->n > 0 && a->a_w != 0.0) {
            const PartState &ps = ctx->ps;
            PartParams pp = ctx->pp;
            hipLaunchKernelGGL((k_interp_field<T>), grid1(a->n, 256), dim3(256), 0, st, pp,
                               (const T *)Rp(a, a->PSI_), ps.p[0], ps.p[1], ps.p[2], a->pv);
            const double dx = ctx->p.lx / ctx->p.nx, dy = ctx->p.ly / ctx->p.ny, dz = ctx->p.lz / ctx->p.nz;
            hipLaunchKernelGGL(k_theta_acc, grid1(a->n, 256), dim3(256), 0, st, a->gth, (const double *)a->theta,
                               (const double *)a->pv, a->n, a->a_w / (dx * dy * dz));
        }
    }

    // <b, dD c> / (2 dt or dt): K-space weighted dot of b and c (reals)
    static void dterm(fipi_adjoint *a, const T *b, const T *c, double *slot)
    {
        fipi_ctx *ctx = a->ctx;
        fwd(ctx, b, Hp(a, 0));
        fwd(ctx, c, Hp(a, 1));
        hipLaunchKernelGGL((k_spec_dot_k4<T>), dim3(DIAG_BLOCKS), dim3(256), 0, ctx->stream, (const C *)Hp(a, 0),
                           (const C *)Hp(a, 1), ctx->d, ctx->kx, ctx->ky, ctx->kz, a->part);
        hipLaunchKernelGGL(k_sum_strided, dim3(1), dim3(256), 0, ctx->stream, (const double *)a->part, DIAG_BLOCKS, 1,
                           1, slot);
    }
};

int adj_snapshot(fipi_adjoint *a, int k)
{
    fipi_ctx *ctx = a->ctx;
    auto bufs = state_buffers(ctx);
    for (size_t i = 0; i < bufs.size(); i++)
        ADJCHK(a, hipMemcpyAsync(a->ckpt[k][i], bufs[i].first, bufs[i].second, hipMemcpyDeviceToDevice, ctx->stream));
    a->ck_t[k] = ctx->t;
    a->ck_cur[k] = ctx->cur;
    a->ck_ready[k] = ctx->conv_ready;
    return FIPI_OK;
}

int adj_restore(fipi_adjoint *a, int k)
{
    fipi_ctx *ctx = a->ctx;
    auto bufs = state_buffers(ctx);
    for (size_t i = 0; i < bufs.size(); i++)
        ADJCHK(a, hipMemcpyAsync(bufs[i].first, a->ckpt[k][i], bufs[i].second, hipMemcpyDeviceToDevice, ctx->stream));
    ctx->t = a->ck_t[k];
    ctx->cur = a->ck_cur[k];
    ctx->conv_ready = a->ck_ready[k];
    return FIPI_OK;
}

// set parameters, wetting field and the initial state (t = 0)
int adj_start(fipi_adjoint *a, double mobility, double lambda, const double *theta)
{
    fipi_ctx *ctx = a->ctx;
    ctx->p.mobility = mobility;
    ctx->p.lambda = lambda;
    if (a->n > 0) {
        ADJCHK(a, hipMemcpyAsync(a->theta, theta, sizeof(double) * a->n, hipMemcpyHostToDevice, ctx->stream));
        // positions first: the wetting source is spread from them
        const PartState &ps = ctx->ps;
        hipLaunchKernelGGL(k_aos_to_soa, grid1(a->n, 256), dim3(256), 0, ctx->stream, (const double *)a->dpos, ps.p[0],
                           ps.p[1], ps.p[2], a->n);
    }
    if (ctx->fp64) Adj<double>::wetting(a);
    else Adj<float>::wetting(a);
    ADJCHK(a, hipStreamSynchronize(ctx->stream));
    ctx->W = a->W;
    // positions and contact forces survive set_state's particle reset: the
    // contact forces are recomputed below
    int rc = fipi_set_state(ctx, a->phi0.data(), a->n ? a->pos0.data() : nullptr, nullptr, nullptr);
    if (rc) return adj_err(a, rc, ctx->err);
    if (a->n > 0) {
        // contact forces of the fixed particles (full ramp), spread every step
        if (ctx->fp64) Impl<double>::cells_build(ctx, ctx->pp, ctx->ps, ctx->cell_ids, ctx->cell_start, ctx->ymax, ctx->yidx);
        else Impl<float>::cells_build(ctx, ctx->pp, ctx->ps, ctx->cell_ids, ctx->cell_start, ctx->ymax, ctx->yidx);
        launch_contact(ctx->stream, a->n, ctx->pp, ctx->ps, ctx->cell_start, ctx->cell_ids, 1.0, (const int *)nullptr,
                       0.0, ctx->d_flags + 1);
        if (ctx->fp64) Adj<double>::contact_field(a);
        else Adj<float>::contact_field(a);
    }
    if ((rc = check_flags(ctx))) return adj_err(a, rc, ctx->err);
    return FIPI_OK;
}

int adj_loss(fipi_adjoint *a, double *loss)
{
    fipi_ctx *ctx = a->ctx;
    if (ctx->fp64) Adj<double>::dot(a, (const double *)ctx->c, (const double *)a->target, 1, a->acc);
    else Adj<float>::dot(a, (const float *)ctx->c, (const float *)a->target, 1, a->acc);
    ADJCHK(a, hipMemcpyAsync(loss, a->acc, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ADJCHK(a, hipStreamSynchronize(ctx->stream));
    return FIPI_OK;
}

template <typename T>
int adj_gradient(fipi_adjoint *a, double *loss, double *g_m, double *g_l, double *g_th)
{
    using A = Adj<T>;
    using C = cplx<T>;
    fipi_ctx *ctx = a->ctx;
    hipStream_t st = ctx->stream;
    const int Tn = a->steps, K = a->every;
    const size_t G = ctx->G;
    int rc;
    // forward with checkpoints at n = 0, K, 2K, ...
    for (int n = 0; n < Tn; n++) {
        if (n % K == 0 && (rc = adj_snapshot(a, n / K))) return rc;
        Impl<T>::step(ctx);
    }
    if ((rc = check_flags(ctx))) return adj_err(a, rc, ctx->err);
    if ((rc = adj_loss(a, loss))) return rc;
    // abar^T = 2 (c^T - c*)
    T *B1 = A::Rp(a, a->B1_), *B2 = A::Rp(a, a->B2_), *B0 = A::Rp(a, a->B0_), *AB = A::Rp(a, a->AB_);
    hipLaunchKernelGGL((k_axpbypc<T>), A::rg(ctx), dim3(256), 0, st, AB, (const T *)ctx->c, 2.0, (const T *)a->target,
                       -2.0, (const T *)nullptr, 0.0, G);
    ADJCHK(a, hipMemsetAsync(a->gth, 0, sizeof(double) * (a->n > 0 ? a->n : 1), st));
    ADJCHK(a, hipMemsetAsync(a->acc, 0, sizeof(double) * 5 * (size_t)(Tn + 1), st));
    ADJCHK(a, hipMemsetAsync(B2, 0, G * sizeof(T), st));
    const double dt = ctx->p.dt, M = ctx->p.mobility, lam = ctx->p.lambda;
    // b for the last step: D^-1 abar^T, or D0^-1 when T = 1
    auto apply_dinv = [&](const T *in, T *out, bool first) {
        C *h = A::Hp(a, 0);
        A::fwd(ctx, in, h);
        hipLaunchKernelGGL((k_spec_scale<T>), A::hg(ctx), dim3(256), 0, st, h, ctx->d, ctx->kx, ctx->ky, ctx->kz,
                           first ? 2 : 1, (first ? 1.0 : 2.0) * dt * M * lam, 1.0 / (double)G);
        A::inv(ctx, h, out);
    };
    if (Tn == 1) apply_dinv(AB, B0, true);
    else apply_dinv(AB, B1, false);
    const int nseg = (Tn + K - 1) / K;
    for (int sgi = nseg - 1; sgi >= 0; sgi--) {
        const int n0 = sgi * K, n1 = std::min(n0 + K, Tn);
        if ((rc = adj_restore(a, sgi))) return rc;
        ADJCHK(a, hipMemcpyAsync(a->seg[0], ctx->c, G * sizeof(T), hipMemcpyDeviceToDevice, st));
        for (int n = n0; n < n1; n++) {
            Impl<T>::step(ctx);
            ADJCHK(a, hipMemcpyAsync(a->seg[n + 1 - n0], ctx->c, G * sizeof(T), hipMemcpyDeviceToDevice, st));
        }
        for (int n = n1 - 1; n >= n0; n--) {
            const T *cn = (const T *)a->seg[n - n0], *cn1 = (const T *)a->seg[n + 1 - n0];
            double *acc_n = a->acc + 5 * (size_t)n;
            // D term of step m = n + 1
            if (n >= 1) A::dterm(a, B1, cn1, acc_n + 3);
            else A::dterm(a, B0, cn1, acc_n + 3);
            // r^n
            if (n >= 1)
                hipLaunchKernelGGL((k_axpbypc<T>), A::rg(ctx), dim3(256), 0, st, A::Rp(a, a->R_), (const T *)B1, 4.0,
                                   (const T *)B2, -2.0, (const T *)nullptr, 0.0, G);
            else
                hipLaunchKernelGGL((k_axpbypc<T>), A::rg(ctx), dim3(256), 0, st, A::Rp(a, a->R_), (const T *)B0, 1.0,
                                   (const T *)B2, -2.0, (const T *)nullptr, 0.0, G);
            ADJCHK(a, hipMemcpyAsync(A::Rp(a, a->C_), cn, G * sizeof(T), hipMemcpyDeviceToDevice, st));
            A::adjoint_at(a, n >= 1, acc_n);
            if (n >= 1) {
                // abar^n = 4 b^{n+1} - b^{n+2} + J_G^T r; then b^n (b0 for n = 1)
                hipLaunchKernelGGL((k_axpbypc<T>), A::rg(ctx), dim3(256), 0, st, AB, (const T *)B1, 4.0, (const T *)B2,
                                   -1.0, (const T *)A::Rp(a, a->JG_), 1.0, G);
                if (n >= 2) {
                    // shift: b^{n+2} <- b^{n+1}, b^{n+1} <- D^-1 abar^n
                    ADJCHK(a, hipMemcpyAsync(B2, B1, G * sizeof(T), hipMemcpyDeviceToDevice, st));
                    apply_dinv(AB, B1, false);
                } else {
                    // n = 1: b0 = D0^-1 abar^1; b^2 stays in B1 -> move to B2
                    ADJCHK(a, hipMemcpyAsync(B2, B1, G * sizeof(T), hipMemcpyDeviceToDevice, st));
                    apply_dinv(AB, B0, true);
                }
            }
        }
    }
    // T = 1: b^2 = 0 and r^0 = b0 (B2 zero) -- handled above since B2 was zeroed
    std::vector<double> acc(5 * (size_t)(Tn + 1));
    ADJCHK(a, hipMemcpyAsync(acc.data(), a->acc, sizeof(double) * acc.size(), hipMemcpyDeviceToHost, st));
    if (a->n > 0) ADJCHK(a, hipMemcpyAsync(g_th, a->gth, sizeof(double) * a->n, hipMemcpyDeviceToHost, st));
    ADJCHK(a, hipStreamSynchronize(st));
    double gm = 0.0, gl = 0.0;
    const double Gd = (double)G;
    for (int n = 0; n < Tn; n++) {
        const double *v = acc.data() + 5 * (size_t)n;
        // <r, -dt L f>;  <r, -dt M L f0> - dt <q, L c + f0>
        gm += -dt * v[0];
        gl += -dt * M * v[1] - dt * v[2];
        // -<b^{n+1}, dD c^{n+1}>: dD/dM = 2 dt lambda k^4 (dt lambda at the first step), dD/dlambda likewise with M
        const double f = n >= 1 ? 2.0 : 1.0;
        gm -= f * dt * lam * v[3] / Gd;
        gl -= f * dt * M * v[3] / Gd;
    }
    *g_m = gm;
    *g_l = gl;
    return check_flags(ctx) ? adj_err(a, FIPI_ERR_RANGE, ctx->err) : FIPI_OK;
}

}  // namespace

extern "C" {

int fipi_adjoint_create(const fipi_params *params, const double *phi0, const double *pos, double wetting,
                        int32_t steps, int32_t checkpoint_every, fipi_adjoint **out)
{
    if (!out) return adj_err(nullptr, FIPI_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!params || !phi0 || steps < 1 || checkpoint_every < 1)
        return adj_err(nullptr, FIPI_ERR_ARG, "params, phi0, steps >= 1 and checkpoint_every >= 1 are required");
    if (params->rho1 != params->rho2)
        return adj_err(nullptr, FIPI_ERR_UNSUPPORTED, "the adjoint needs rho1 == rho2 (no buoyancy contrast)");
    if (params->interp_u != 0 || params->kt != 0.0)
        return adj_err(nullptr, FIPI_ERR_UNSUPPORTED, "the adjoint's particles are fixed: interp_u and kt must be 0");
    const int n = params->switch_particle ? params->n_particles : 0;
    if (n > 0 && !pos) return adj_err(nullptr, FIPI_ERR_ARG, "pos is required");
    auto *a = new fipi_adjoint();
    a->steps = steps;
    a->every = checkpoint_every;
    a->a_w = wetting;
    a->n = n;
    int rc = fipi_create(params, &a->ctx);
    if (rc) {
        delete a;
        return adj_err(nullptr, rc, g_last_error);
    }
    fipi_ctx *ctx = a->ctx;
    ctx->fix_particles = true;
    // the checkpoints copy the real velocity: keep it written every step
    ctx->u_lazy = false;
    const size_t G = ctx->G;
    a->phi0.assign(phi0, phi0 + G);
    if (n > 0) a->pos0.assign(pos, pos + 3 * (size_t)n);
    const size_t R = G * ctx->esz, H = ctx->GH * 2 * ctx->esz;
#define AALLOC(p, b) \
    if ((rc = adj_alloc(a, (void **)&(p), (b))) != FIPI_OK) { fipi_adjoint_destroy(a); return rc; }
    AALLOC(a->target, R);
    AALLOC(a->W, R);
    AALLOC(a->spread3, 3 * R);
    AALLOC(a->theta, sizeof(double) * (n + 1));
    AALLOC(a->dpos, sizeof(double) * 3 * (n + 1));
    AALLOC(a->wgt, sizeof(double) * (n + 1));
    AALLOC(a->zero, sizeof(double) * (n + 1));
    AALLOC(a->gth, sizeof(double) * (n + 1));
    AALLOC(a->pv, sizeof(double) * (n + 1));
    AALLOC(a->part, sizeof(double) * (DIAG_BLOCKS + 1));
    AALLOC(a->acc, sizeof(double) * 5 * (size_t)(steps + 1));
    for (int k = 0; k < fipi_adjoint::NREAL; k++) AALLOC(a->rw[k], R);
    for (int k = 0; k < 4; k++) AALLOC(a->hw[k], H);
    const int nck = (steps + checkpoint_every - 1) / checkpoint_every;
    a->ckpt.resize(nck);
    a->ck_t.resize(nck);
    a->ck_cur.resize(nck);
    a->ck_ready.resize(nck);
    auto bufs = state_buffers(ctx);
    for (int k = 0; k < nck; k++)
        for (auto &b : bufs) {
            void *p;
            AALLOC(p, b.second);
            a->ckpt[k].push_back(p);
        }
    a->seg.resize(checkpoint_every + 1);
    for (auto &p : a->seg) AALLOC(p, R);
#undef AALLOC
    hipError_t he = hipMemset(a->zero, 0, sizeof(double) * (n + 1));
    for (int k = 0; k < fipi_adjoint::NREAL && he == hipSuccess; k++) he = hipMemset(a->rw[k], 0, R);
    if (he == hipSuccess && n > 0) he = hipMemcpy(a->dpos, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        fipi_adjoint_destroy(a);
        return adj_err(nullptr, FIPI_ERR_HIP, hipGetErrorString(he));
    }
    *out = a;
    return FIPI_OK;
}

int fipi_adjoint_destroy(fipi_adjoint *a)
{
    if (!a) return FIPI_OK;
    if (a->ctx) {
        (void)hipStreamSynchronize(a->ctx->stream);
        a->ctx->W = nullptr;
    }
    for (void *p : a->mem) (void)hipFree(p);
    fipi_destroy(a->ctx);
    delete a;
    return FIPI_OK;
}

const char *fipi_adjoint_last_error(const fipi_adjoint *a) { return a ? a->err.c_str() : g_last_error.c_str(); }

int fipi_adjoint_set_target(fipi_adjoint *a, const double *phi_target)
{
    if (!a || !phi_target) return adj_err(a, FIPI_ERR_ARG, "invalid argument");
    int rc = copy_in_real(a->ctx, a->target, phi_target, a->ctx->G);
    if (rc) return adj_err(a, rc, a->ctx->err);
    ADJCHK(a, hipStreamSynchronize(a->ctx->stream));
    a->has_target = true;
    return FIPI_OK;
}

int fipi_adjoint_forward(fipi_adjoint *a, double mobility, double lambda, const double *theta, double *loss,
                         double *phi_final)
{
    if (!a || (a->n > 0 && !theta)) return adj_err(a, FIPI_ERR_ARG, "invalid argument");
    int rc = adj_start(a, mobility, lambda, theta);
    if (rc) return rc;
    fipi_ctx *ctx = a->ctx;
    for (int n = 0; n < a->steps; n++) {
        if (ctx->fp64) Impl<double>::step(ctx);
        else Impl<float>::step(ctx);
    }
    if ((rc = check_flags(ctx))) return adj_err(a, rc, ctx->err);
    if (phi_final && (rc = copy_out_real(ctx, phi_final, ctx->c, ctx->G))) return adj_err(a, rc, ctx->err);
    if (loss) {
        if (!a->has_target) return adj_err(a, FIPI_ERR_STATE, "no target set");
        if ((rc = adj_loss(a, loss))) return rc;
    }
    return FIPI_OK;
}

int fipi_adjoint_gradient(fipi_adjoint *a, double mobility, double lambda, const double *theta, double *loss,
                          double *g_mobility, double *g_lambda, double *g_theta)
{
    if (!a || !loss || !g_mobility || !g_lambda || (a->n > 0 && (!theta || !g_theta)))
        return adj_err(a, FIPI_ERR_ARG, "invalid argument");
    if (!a->has_target) return adj_err(a, FIPI_ERR_STATE, "no target set");
    int rc = adj_start(a, mobility, lambda, theta);
    if (rc) return rc;
    return a->ctx->fp64 ? adj_gradient<double>(a, loss, g_mobility, g_lambda, g_theta)
                        : adj_gradient<float>(a, loss, g_mobility, g_lambda, g_theta);
}

}  // extern "C"
