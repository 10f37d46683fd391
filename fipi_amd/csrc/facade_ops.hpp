// Device-side operations of the class facade (fipi_amd/facade/fipi.hpp):
// buffers, the complex 3-D transform, the `field` algebra and the
// per-particle entries that main.cpp's loop body reaches through the
// `cluster` methods.  Included at the end of fipi.hip.
//
// A facade field is `dim` components of G complex values (interleaved
// re, im) in device memory, in real space or in K space, as the
// reference's field holds fftw_complex arrays (src/field.hpp:35-93).  The
// reference's operators loop over every point; here each is one
// elementwise kernel with the same arithmetic per point.
#pragma once

namespace {

// kernels ------------------------------------------------------------------

__global__ void k_cf_split(const double2 *__restrict__ v, double *__restrict__ re, double *__restrict__ im, size_t G)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < G) {
        re[i] = v[i].x;
        im[i] = v[i].y;
    }
}

__global__ void k_cf_join(const double *__restrict__ re, const double *__restrict__ im, double2 *__restrict__ v, size_t G)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < G) v[i] = double2{re[i], im[i]};
}

// X = FFT(re) + i FFT(im) from the two half spectra (kx <= nx/2) of the
// real transforms; the other half from Hermitian symmetry
__global__ void k_cf_expand(const double2 *__restrict__ h1, const double2 *__restrict__ h2, double2 *__restrict__ v, int nx,
                            int ny, int nz)
{
    const size_t G = (size_t)nx * ny * nz;
    const size_t it = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= G) return;
    const int nxh = nx / 2 + 1;
    const int i = (int)(it % nx), j = (int)((it / nx) % ny), k = (int)(it / ((size_t)nx * ny));
    if (i <= nx / 2) {
        const size_t q = (size_t)i + (size_t)nxh * (j + (size_t)ny * k);
        const double2 a = h1[q], b = h2[q];
        v[it] = double2{a.x - b.y, a.y + b.x};
    } else {
        const int im = nx - i, jm = (ny - j) % ny, km = (nz - k) % nz;
        const size_t q = (size_t)im + (size_t)nxh * (jm + (size_t)ny * km);
        const double2 a = h1[q], b = h2[q];
        v[it] = double2{a.x + b.y, b.x - a.y};
    }
}

// Hermitian part A = (X(k) + conj X(-k))/2 and -i B, B = (X(k) - conj
// X(-k))/2, on the half spectrum: the real transforms of A and -iB are the
// real and imaginary parts of the inverse transform of X
__global__ void k_cf_herm(const double2 *__restrict__ v, double2 *__restrict__ ha, double2 *__restrict__ hb, int nx, int ny,
                          int nz)
{
    const int nxh = nx / 2 + 1;
    const size_t n = (size_t)nxh * ny * nz;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = (int)(q % nxh), j = (int)((q / nxh) % ny), k = (int)(q / ((size_t)nxh * ny));
    const int im = (nx - i) % nx, jm = (ny - j) % ny, km = (nz - k) % nz;
    const double2 x = v[(size_t)i + (size_t)nx * (j + (size_t)ny * k)];
    const double2 y = v[(size_t)im + (size_t)nx * (jm + (size_t)ny * km)];
    ha[q] = double2{(x.x + y.x) * 0.5, (x.y - y.y) * 0.5};
    const double br = (x.x - y.x) * 0.5, bi = (x.y + y.y) * 0.5;
    hb[q] = double2{bi, -br};
}

// The field algebra, one point per thread.  Operation codes and semantics
// follow src/field.cpp (cited per case); a_k / b_k: the operand is in K space.
__global__ void k_field_op(int op, double2 *__restrict__ out, int out_dim, const double2 *__restrict__ a, int a_dim,
                           int a_k, const double2 *__restrict__ b, int b_dim, int b_k, double s, int nx, int ny, int nz)
{
    const size_t G = (size_t)nx * ny * nz;
    const size_t it = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= G) return;
    const bool k_any = a_k || b_k;
    switch (op) {
    case FIPI_OP_ADD:  // field.cpp:525-570, out is the left operand
    case FIPI_OP_SUB:  // field.cpp:598-640
        for (int d = 0; d < out_dim; d++) {
            double2 x = out[d * G + it];
            const double2 y = b[d * G + it];
            if (op == FIPI_OP_ADD) x.x += y.x;
            else x.x -= y.x;
            if (a_k) x.y = op == FIPI_OP_ADD ? x.y + y.y : x.y - y.y;
            else x.y = 0.0;
            out[d * G + it] = x;
        }
        break;
    case FIPI_OP_ADD_S:  // field.cpp:574-593 (real space only)
    case FIPI_OP_SUB_S:  // field.cpp:644-662
        for (int d = 0; d < out_dim; d++) {
            double2 x = out[d * G + it];
            x.x = op == FIPI_OP_ADD_S ? x.x + s : x.x - s;
            x.y = 0.0;
            out[d * G + it] = x;
        }
        break;
    case FIPI_OP_MUL_S:  // field.cpp:711-738
        for (int d = 0; d < out_dim; d++) {
            double2 x = out[d * G + it];
            x.x *= s;
            x.y = a_k ? x.y * s : 0.0;
            out[d * G + it] = x;
        }
        break;
    case FIPI_OP_DIV_S:  // field.cpp:790-830, division by zero gives zero
        for (int d = 0; d < out_dim; d++) {
            double2 x = out[d * G + it];
            if (s != 0.0) {
                x.x /= s;
                x.y = a_k ? x.y / s : 0.0;
            } else {
                x.x = 0.0;
                x.y = 0.0;
            }
            out[d * G + it] = x;
        }
        break;
    case FIPI_OP_MULEQ: {  // field.cpp:666-707, b a scalar field
        const double2 r = b[it];
        for (int d = 0; d < out_dim; d++) {
            double2 x = out[d * G + it];
            if (k_any) {
                const double cr = -x.y * r.y, ci = x.x * r.y;
                x.x *= r.x;
                x.x += cr;
                x.y *= r.x;
                x.y += ci;
            } else {
                x.x *= r.x;
                x.y = 0.0;
            }
            out[d * G + it] = x;
        }
        break;
    }
    case FIPI_OP_DIVEQ: {  // field.cpp:742-786, b a real-space scalar field
        const double r = b[it].x;
        for (int d = 0; d < out_dim; d++) {
            double2 x = out[d * G + it];
            if (r != 0.0) {
                x.x /= r;
                x.y = a_k ? x.y / r : 0.0;
            } else {
                x.x = 0.0;
                x.y = 0.0;
            }
            out[d * G + it] = x;
        }
        break;
    }
    case FIPI_OP_MUL:  // field.cpp:946-990: out[(d1+1)(d2+1)-1] = a[d1] * b[d2]
    case FIPI_OP_DIV:  // field.cpp:1002-1046, b real space, out in a's space
        for (int d1 = 0; d1 < a_dim; d1++)
            for (int d2 = 0; d2 < b_dim; d2++) {
                const double2 x = a[d1 * G + it], y = b[d2 * G + it];
                double2 o;
                if (op == FIPI_OP_MUL) {
                    if (k_any) o = double2{x.x * y.x - x.y * y.y, x.y * y.x + x.x * y.y};
                    else o = double2{x.x * y.x, 0.0};
                } else {
                    if (a_k) o = y.x != 0.0 ? double2{x.x / y.x, x.y / y.x} : double2{0.0, 0.0};
                    else o = double2{y.x != 0.0 ? x.x / y.x : 0.0, 0.0};
                }
                out[((d1 + 1) * (d2 + 1) - 1) * G + it] = o;
            }
        break;
    case FIPI_OP_DOT: {  // field.cpp:364-405
        double2 o{0.0, 0.0};
        for (int d = 0; d < a_dim; d++) {
            const double2 x = a[d * G + it], y = b[d * G + it];
            if (k_any) {
                o.x += x.x * y.x - x.y * y.y;
                o.y += x.y * y.x + x.x * y.y;
            } else {
                o.x += x.x * y.x;
            }
        }
        if (!k_any) o.y = 0.0;
        out[it] = o;
        break;
    }
    case FIPI_OP_POWER: {  // field.cpp:408-431: sum over components of pow(re, s)
        double o = 0.0;
        for (int d = 0; d < a_dim; d++) o += pow(a[d * G + it].x, s);
        out[it] = double2{o, 0.0};
        break;
    }
    case FIPI_OP_GRAD: {  // field.cpp:288-328: i k_d a, zero on the Nyquist plane of axis d
        const int i = (int)(it % nx), j = (int)((it / nx) % ny), k = (int)(it / ((size_t)nx * ny));
        const double2 x = a[it];
        for (int d = 0; d < 3; d++) {
            const bool nyq = (d == 0 && i == nx / 2) || (d == 1 && j == ny / 2) || (d == 2 && k == nz / 2);
            const double kv = b[d * G + it].x;
            out[d * G + it] = nyq ? double2{0.0, 0.0} : double2{-kv * x.y, kv * x.x};
        }
        break;
    }
    case FIPI_OP_COPY:  // field.cpp:836-866
        for (int d = 0; d < out_dim; d++) out[d * G + it] = a[d * G + it];
        break;
    case FIPI_OP_RESET:  // field.cpp:179-193
        for (int d = 0; d < out_dim; d++) out[d * G + it] = double2{0.0, 0.0};
        break;
    case FIPI_OP_REAL: {  // real parts into a real array (out as doubles)
        double *o = reinterpret_cast<double *>(out);
        for (int d = 0; d < a_dim; d++) o[d * G + it] = a[d * G + it].x;
        break;
    }
    case FIPI_OP_FROM_REAL: {  // a real array (a as doubles) into a real-space field
        const double *r = reinterpret_cast<const double *>(a);
        for (int d = 0; d < out_dim; d++) out[d * G + it] = double2{r[d * G + it], 0.0};
        break;
    }
    default:
        break;
    }
}

// element::get_f_pi for n particles (AoS vectors)
__global__ void k_f_pi_unit(PartParams P, const double *__restrict__ cloc, const double *__restrict__ dcdx,
                            const double *__restrict__ rad, const double *__restrict__ cth, double *__restrict__ fpi)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.n) return;
    double f[3];
    capillary_force(P, cloc[a], dcdx[3 * a], dcdx[3 * a + 1], dcdx[3 * a + 2], rad ? rad[a] : P.r_p,
                    cth ? cth[a] : 0.0, f);
    fpi[3 * a] = f[0];
    fpi[3 * a + 1] = f[1];
    fpi[3 * a + 2] = f[2];
}

// element::advect (stage 1 or 2) or element::enforce_bc (stage 0), AoS
__global__ void k_advect_unit(PartParams P, int stage, double *__restrict__ p, double *__restrict__ po,
                              double *__restrict__ u, double *__restrict__ uo, const double *__restrict__ fpi,
                              const double *__restrict__ fpp, const double *__restrict__ rad)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.n) return;
    const double L[3] = {P.lx, P.ly, P.lz};
    for (int d = 0; d < 3; d++) {
        const size_t q = 3 * (size_t)a + d;
        if (stage == 0) {
            p[q] = fmod(p[q] + L[d], L[d]);
            continue;
        }
        const double v = heun_velocity(P, fpi[q], 0.0, fpp[q], rad ? rad[a] : P.r_p, 0.0);
        if (stage == 1) {
            uo[q] = v;
            po[q] = p[q];
            p[q] = p[q] + v * P.dt;
        } else {
            u[q] = v;
            p[q] = po[q] + (uo[q] + v) * (0.5 * P.dt);
        }
    }
}

// facade scratch of a context: one real field and two half spectra
struct FacadeScratch {
    double *re = nullptr, *im = nullptr;
    double2 *h1 = nullptr, *h2 = nullptr;
};

}  // namespace

static int facade_scratch(fipi_ctx *ctx, FacadeScratch *fs)
{
    const size_t G = ctx->G, GHu = (size_t)ctx->d.nxh * ctx->p.ny * ctx->p.nz;
    if (!ctx->fs_re) {
        int rc;
        if ((rc = dalloc(ctx, &ctx->fs_re, sizeof(double) * G))) return rc;
        if ((rc = dalloc(ctx, &ctx->fs_im, sizeof(double) * G))) return rc;
        if ((rc = dalloc(ctx, &ctx->fs_h1, sizeof(double2) * GHu))) return rc;
        if ((rc = dalloc(ctx, &ctx->fs_h2, sizeof(double2) * GHu))) return rc;
    }
    fs->re = (double *)ctx->fs_re;
    fs->im = (double *)ctx->fs_im;
    fs->h1 = (double2 *)ctx->fs_h1;
    fs->h2 = (double2 *)ctx->fs_h2;
    return FIPI_OK;
}

extern "C" {

int fipi_buffer_create(fipi_ctx *ctx, int64_t bytes, void **d_ptr)
{
    if (!ctx || !d_ptr || bytes < 0) return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    *d_ptr = nullptr;
    void *p = nullptr;
    HIPCHK(ctx, hipMalloc(&p, bytes > 0 ? (size_t)bytes : 16));
    const hipError_t e = hipMemsetAsync(p, 0, bytes > 0 ? (size_t)bytes : 16, ctx->stream);
    if (e == hipSuccess) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        (void)hipFree(p);
        return set_err(ctx, FIPI_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    }
    *d_ptr = p;
    return FIPI_OK;
}

int fipi_buffer_destroy(fipi_ctx *ctx, void *d_ptr)
{
    if (!ctx) return set_err(ctx, FIPI_ERR_ARG, "ctx is NULL");
    if (!d_ptr) return FIPI_OK;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipFree(d_ptr));
    return FIPI_OK;
}

int fipi_buffer_upload(fipi_ctx *ctx, void *d_dst, const void *h_src, int64_t bytes)
{
    if (!ctx || bytes < 0 || (bytes > 0 && (!d_dst || !h_src))) return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    if (bytes == 0) return FIPI_OK;
    HIPCHK(ctx, hipMemcpyAsync(d_dst, h_src, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return FIPI_OK;
}

int fipi_buffer_download(fipi_ctx *ctx, void *h_dst, const void *d_src, int64_t bytes)
{
    if (!ctx || bytes < 0 || (bytes > 0 && (!d_src || !h_dst))) return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    if (bytes == 0) return FIPI_OK;
    HIPCHK(ctx, hipMemcpyAsync(h_dst, d_src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return FIPI_OK;
}

int fipi_field_fft(fipi_ctx *ctx, double *d_v, int32_t dim, int32_t sign)
{
    if (!ctx || !d_v || dim < 1 || (sign != -1 && sign != 1)) return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    if (!ctx->fp64 || ctx->nbox != 1)
        return set_err(ctx, FIPI_ERR_UNSUPPORTED, "the facade's complex transform needs a single-box FP64 context");
    FacadeScratch fs;
    int rc = facade_scratch(ctx, &fs);
    if (rc) return rc;
    const size_t G = ctx->G, GHu = (size_t)ctx->d.nxh * ctx->p.ny * ctx->p.nz;
    const int nx = ctx->p.nx, ny = ctx->p.ny, nz = ctx->p.nz;
    hipStream_t st = ctx->stream;
    for (int d = 0; d < dim; d++) {
        double2 *v = reinterpret_cast<double2 *>(d_v) + (size_t)d * G;
        if (sign < 0) {
            // FFT(re + i im) = FFT(re) + i FFT(im), unnormalised (FFTW_FORWARD)
            hipLaunchKernelGGL(k_cf_split, grid1(G, 256), dim3(256), 0, st, (const double2 *)v, fs.re, fs.im, G);
            if ((rc = fipi_fft_r2c(ctx, fs.re, fs.h1))) return rc;
            if ((rc = fipi_fft_r2c(ctx, fs.im, fs.h2))) return rc;
            hipLaunchKernelGGL(k_cf_expand, grid1(G, 256), dim3(256), 0, st, (const double2 *)fs.h1,
                               (const double2 *)fs.h2, v, nx, ny, nz);
        } else {
            // inverse, normalised by 1/G (field::ifft + normalize)
            hipLaunchKernelGGL(k_cf_herm, grid1(GHu, 256), dim3(256), 0, st, (const double2 *)v, fs.h1, fs.h2, nx, ny,
                               nz);
            if ((rc = fipi_fft_c2r(ctx, fs.h1, fs.re))) return rc;
            if ((rc = fipi_fft_c2r(ctx, fs.h2, fs.im))) return rc;
            hipLaunchKernelGGL(k_cf_join, grid1(G, 256), dim3(256), 0, st, (const double *)fs.re,
                               (const double *)fs.im, v, G);
        }
    }
    HIPCHK(ctx, hipStreamSynchronize(st));
    HIPCHK(ctx, hipGetLastError());
    return FIPI_OK;
}

int fipi_field_op(fipi_ctx *ctx, int32_t op, double *d_out, int32_t out_dim, const double *d_a, int32_t a_dim,
                  int32_t a_kspace, const double *d_b, int32_t b_dim, int32_t b_kspace, double s)
{
    if (!ctx || !d_out || op < FIPI_OP_ADD || op > FIPI_OP_FROM_REAL) return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    if (ctx->nbox != 1) return set_err(ctx, FIPI_ERR_UNSUPPORTED, "field algebra needs a single-box context");
    auto bad = [&](bool c) { return c; };
    if (out_dim < 1 || out_dim > 9 || (d_a && (a_dim < 1 || a_dim > 9)) || (d_b && (b_dim < 1 || b_dim > 9)))
        return set_err(ctx, FIPI_ERR_ARG, "field dimensions are 1, 3 or 9");
    const bool needs_a = op == FIPI_OP_MUL || op == FIPI_OP_DIV || op == FIPI_OP_DOT || op == FIPI_OP_POWER ||
                         op == FIPI_OP_GRAD || op == FIPI_OP_COPY || op == FIPI_OP_REAL || op == FIPI_OP_FROM_REAL;
    const bool needs_b = op == FIPI_OP_ADD || op == FIPI_OP_SUB || op == FIPI_OP_MULEQ || op == FIPI_OP_DIVEQ ||
                         op == FIPI_OP_MUL || op == FIPI_OP_DIV || op == FIPI_OP_DOT || op == FIPI_OP_GRAD;
    if (bad(needs_a && !d_a) || bad(needs_b && !d_b)) return set_err(ctx, FIPI_ERR_ARG, "missing operand");
    if ((op == FIPI_OP_ADD || op == FIPI_OP_SUB) && (b_dim != out_dim || a_kspace != b_kspace))
        return set_err(ctx, FIPI_ERR_ARG, "+= / -= need two fields of one dimension in one space (field.cpp:529-542)");
    if ((op == FIPI_OP_ADD_S || op == FIPI_OP_SUB_S) && a_kspace)
        return set_err(ctx, FIPI_ERR_ARG, "adding a real number in frequency space (field.cpp:579-583)");
    if ((op == FIPI_OP_MULEQ || op == FIPI_OP_DIVEQ) && b_dim != 1)
        return set_err(ctx, FIPI_ERR_ARG, "*= / /= need a scalar right-hand field (field.cpp:702-706)");
    if ((op == FIPI_OP_DIVEQ || op == FIPI_OP_DIV) && b_kspace)
        return set_err(ctx, FIPI_ERR_ARG, "divisor field in frequency space (field.cpp:745-749)");
    if (op == FIPI_OP_DOT && a_dim != b_dim) return set_err(ctx, FIPI_ERR_ARG, "dot() of fields of different dimensions");
    if (op == FIPI_OP_POWER && a_kspace) return set_err(ctx, FIPI_ERR_ARG, "power() of a field in frequency space");
    if (op == FIPI_OP_GRAD && (!a_kspace || a_dim != 1 || b_dim != 3 || out_dim != 3))
        return set_err(ctx, FIPI_ERR_ARG, "grad() of a scalar field in frequency space with a wave-vector field");
    if ((op == FIPI_OP_MUL || op == FIPI_OP_DIV) && out_dim != a_dim * b_dim)
        return set_err(ctx, FIPI_ERR_ARG, "product field dimension");
    const size_t G = ctx->G;
    hipLaunchKernelGGL(k_field_op, grid1(G, 256), dim3(256), 0, ctx->stream, (int)op, reinterpret_cast<double2 *>(d_out),
                       (int)out_dim, reinterpret_cast<const double2 *>(d_a), (int)a_dim, (int)a_kspace,
                       reinterpret_cast<const double2 *>(d_b), (int)b_dim, (int)b_kspace, s, ctx->p.nx, ctx->p.ny,
                       ctx->p.nz);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    return FIPI_OK;
}

int fipi_capillary_forces(fipi_ctx *ctx, const double *d_c_loc, const double *d_dcdx, const double *d_radius,
                          const double *d_cos_theta, int32_t n, double *d_f_pi)
{
    if (!ctx || n < 0 || (n > 0 && (!d_c_loc || !d_dcdx || !d_f_pi))) return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    if (n == 0) return FIPI_OK;
    PartParams pp = ctx->pp;
    pp.n = n;
    hipLaunchKernelGGL(k_f_pi_unit, grid1(n, 256), dim3(256), 0, ctx->stream, pp, d_c_loc, d_dcdx, d_radius,
                       d_cos_theta, d_f_pi);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    return FIPI_OK;
}

int fipi_advect(fipi_ctx *ctx, int32_t stage, int32_t n, double *d_pos, double *d_pos_old, double *d_u, double *d_u_old,
                const double *d_f_pi, const double *d_f_pp, const double *d_radius)
{
    if (!ctx || n < 0 || stage < 0 || stage > 2 || (n > 0 && !d_pos) ||
        (n > 0 && stage > 0 && (!d_pos_old || !d_u || !d_u_old || !d_f_pi || !d_f_pp)))
        return set_err(ctx, FIPI_ERR_ARG, "invalid argument");
    if (n == 0) return FIPI_OK;
    PartParams pp = ctx->pp;
    pp.n = n;
    hipLaunchKernelGGL(k_advect_unit, grid1(n, 256), dim3(256), 0, ctx->stream, pp, (int)stage, d_pos, d_pos_old, d_u,
                       d_u_old, d_f_pi, d_f_pp, d_radius);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipGetLastError());
    return FIPI_OK;
}

}  // extern "C"
