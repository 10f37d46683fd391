/*
 * CPU FP64 oracle for the FIPI per-timestep hot path -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference time loop body
 * (src/main.cpp:75-222) and the routines it calls.  It is the checker the
 * GPU library is compared with and the CPU baseline timed beside it; the
 * product library never links or loads it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 *
 * Field storage mirrors the reference `field` class (src/field.hpp:21-31):
 * each component is an interleaved complex array of G values, imaginary
 * part carried in real space too, with a per-field "in k-space" flag that
 * this restatement tracks implicitly through the fixed statement order.
 * Every statement of main.cpp is written as one loop whose per-element
 * arithmetic is the composition of the reference operators it calls, in
 * the same association order (operator bodies: src/field.cpp:528-1080).
 */
#include "fipi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ======================================================================
 * 1-D Stockham FFT (stand-in for FFTW's fftw_execute_dft; the DFT
 * definition, not FFTW's code, is what this follows).
 *   X[m] = sum_n x[n] exp(sign * 2*pi*i*n*m/N),  unnormalised.
 * Stockham autosort stages after Govindaraju et al., "High performance
 * discrete Fourier transforms on graphics processors" (SC'08).
 * ====================================================================== */
typedef struct {
    int n;
    int nf;
    int fac[40];
    double *tw;     /* tw[2t], tw[2t+1] = cos, sin (2*pi*t/n), t < n */
} fft_plan1d;

/* Factor order of every plan made after the call: 0 (default) radix-4
 * stages then radix 2, 1 radix-2 stages only.  The two orders round
 * differently (~1e-16 per transform); the long-horizon tests use the pair
 * to show how far two correct FP64 implementations of the step agree. */
static int g_fft_order = 0;

void oracle_set_fft_order(int order) { g_fft_order = order; }

static void plan1d_init(fft_plan1d *pl, int n)
{
    pl->n = n;
    pl->nf = 0;
    int m = n;
    if (g_fft_order == 0)
        while (m % 4 == 0) { pl->fac[pl->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { pl->fac[pl->nf++] = 2; m /= 2; }
    for (int r = 3; m > 1; r += 2)
        while (m % r == 0) { pl->fac[pl->nf++] = r; m /= r; }
    pl->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < n; t++) {
        long double a = two_pi * (long double)t / (long double)n;
        pl->tw[2 * t] = (double)cosl(a);
        pl->tw[2 * t + 1] = (double)sinl(a);
    }
}

static void plan1d_free(fft_plan1d *pl) { free(pl->tw); pl->tw = NULL; }

/* generic small DFT of radix r on v (interleaved), in place, via buffer */
static void dft_small(double *v, int r, int sign, const fft_plan1d *pl)
{
    double out[2 * 64];
    const int step = pl->n / r; /* exp(2 pi i q m / r) = tw[(q*m mod r)*step] */
    for (int m = 0; m < r; m++) {
        double sr = 0.0, si = 0.0;
        for (int q = 0; q < r; q++) {
            int t = (q * m) % r;
            double c = pl->tw[2 * t * step];
            double s = sign * pl->tw[2 * t * step + 1];
            sr += v[2 * q] * c - v[2 * q + 1] * s;
            si += v[2 * q] * s + v[2 * q + 1] * c;
        }
        out[2 * m] = sr;
        out[2 * m + 1] = si;
    }
    memcpy(v, out, sizeof(double) * 2 * r);
}

/* One 1-D transform: x (interleaved, length n) -> x, using scratch y. */
static void fft1d(const fft_plan1d *pl, double *x, double *y, int sign)
{
    const int n = pl->n;
    double *src = x, *dst = y;
    int ns = 1;
    for (int f = 0; f < pl->nf; f++) {
        const int r = pl->fac[f];
        const int nr = n / r;
        for (int j = 0; j < nr; j++) {
            double v[2 * 64];
            const int k = j % ns;
            for (int q = 0; q < r; q++) {
                v[2 * q] = src[2 * (j + q * nr)];
                v[2 * q + 1] = src[2 * (j + q * nr) + 1];
            }
            if (ns > 1) {
                const int tstep = n / (ns * r);
                for (int q = 1; q < r; q++) {
                    int t = (k * q * tstep) % n;
                    double c = pl->tw[2 * t];
                    double s = sign * pl->tw[2 * t + 1];
                    double a = v[2 * q], b = v[2 * q + 1];
                    v[2 * q] = a * c - b * s;
                    v[2 * q + 1] = a * s + b * c;
                }
            }
            if (r == 2) {
                double ar = v[0], ai = v[1], br = v[2], bi = v[3];
                v[0] = ar + br; v[1] = ai + bi;
                v[2] = ar - br; v[3] = ai - bi;
            } else if (r == 4) {
                double ar = v[0], ai = v[1], br = v[2], bi = v[3];
                double cr = v[4], ci = v[5], dr = v[6], di = v[7];
                double t0r = ar + cr, t0i = ai + ci;
                double t1r = ar - cr, t1i = ai - ci;
                double t2r = br + dr, t2i = bi + di;
                double t3r = br - dr, t3i = bi - di;
                /* sign*i*(t3) = (-sign*t3i, sign*t3r) */
                v[0] = t0r + t2r; v[1] = t0i + t2i;
                v[4] = t0r - t2r; v[5] = t0i - t2i;
                v[2] = t1r - sign * t3i; v[3] = t1i + sign * t3r;
                v[6] = t1r + sign * t3i; v[7] = t1i - sign * t3r;
            } else {
                dft_small(v, r, sign, pl);
            }
            const int base = (j / ns) * ns * r + k;
            for (int q = 0; q < r; q++) {
                dst[2 * (base + q * ns)] = v[2 * q];
                dst[2 * (base + q * ns) + 1] = v[2 * q + 1];
            }
        }
        double *t = src; src = dst; dst = t;
        ns *= r;
    }
    if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)n);
}

#define OFFT_BLOCK 8

int oracle_fft3d(int nx, int ny, int nz, double *data, int sign, int n_threads)
{
    if (nx < 1 || ny < 1 || nz < 1) return -1;
    fft_plan1d px, py, pz;
    plan1d_init(&px, nx);
    plan1d_init(&py, ny);
    plan1d_init(&pz, nz);
    const size_t sxy = (size_t)nx * ny;
    int nt = n_threads > 0 ? n_threads : omp_get_max_threads();

    /* x lines: contiguous */
    #pragma omp parallel num_threads(nt)
    {
        double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)nx);
        #pragma omp for schedule(static)
        for (long l = 0; l < (long)ny * nz; l++)
            fft1d(&px, data + 2 * (size_t)l * nx, scr, sign);
        free(scr);
    }
    /* y lines: stride nx; gather OFFT_BLOCK consecutive x at a time */
    #pragma omp parallel num_threads(nt)
    {
        double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)ny * OFFT_BLOCK);
        double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)ny);
        const int nxb = (nx + OFFT_BLOCK - 1) / OFFT_BLOCK;
        #pragma omp for schedule(static)
        for (long w = 0; w < (long)nz * nxb; w++) {
            const int k = (int)(w / nxb), i0 = (int)(w % nxb) * OFFT_BLOCK;
            const int bw = (i0 + OFFT_BLOCK <= nx) ? OFFT_BLOCK : nx - i0;
            double *base = data + 2 * ((size_t)k * sxy + i0);
            for (int j = 0; j < ny; j++)
                for (int b = 0; b < bw; b++) {
                    buf[2 * ((size_t)b * ny + j)] = base[2 * ((size_t)j * nx + b)];
                    buf[2 * ((size_t)b * ny + j) + 1] = base[2 * ((size_t)j * nx + b) + 1];
                }
            for (int b = 0; b < bw; b++) fft1d(&py, buf + 2 * (size_t)b * ny, scr, sign);
            for (int j = 0; j < ny; j++)
                for (int b = 0; b < bw; b++) {
                    base[2 * ((size_t)j * nx + b)] = buf[2 * ((size_t)b * ny + j)];
                    base[2 * ((size_t)j * nx + b) + 1] = buf[2 * ((size_t)b * ny + j) + 1];
                }
        }
        free(buf); free(scr);
    }
    /* z lines: stride nx*ny */
    #pragma omp parallel num_threads(nt)
    {
        double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)nz * OFFT_BLOCK);
        double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)nz);
        const int nxb = (nx + OFFT_BLOCK - 1) / OFFT_BLOCK;
        #pragma omp for schedule(static)
        for (long w = 0; w < (long)ny * nxb; w++) {
            const int j = (int)(w / nxb), i0 = (int)(w % nxb) * OFFT_BLOCK;
            const int bw = (i0 + OFFT_BLOCK <= nx) ? OFFT_BLOCK : nx - i0;
            double *base = data + 2 * ((size_t)j * nx + i0);
            for (int k = 0; k < nz; k++)
                for (int b = 0; b < bw; b++) {
                    buf[2 * ((size_t)b * nz + k)] = base[2 * ((size_t)k * sxy + b)];
                    buf[2 * ((size_t)b * nz + k) + 1] = base[2 * ((size_t)k * sxy + b) + 1];
                }
            for (int b = 0; b < bw; b++) fft1d(&pz, buf + 2 * (size_t)b * nz, scr, sign);
            for (int k = 0; k < nz; k++)
                for (int b = 0; b < bw; b++) {
                    base[2 * ((size_t)k * sxy + b)] = buf[2 * ((size_t)b * nz + k)];
                    base[2 * ((size_t)k * sxy + b) + 1] = buf[2 * ((size_t)b * nz + k) + 1];
                }
        }
        free(buf); free(scr);
    }
    plan1d_free(&px); plan1d_free(&py); plan1d_free(&pz);
    return 0;
}

/* ======================================================================
 * Helpers restating src/primitive.{hpp,cpp}
 * ====================================================================== */

/* vector::norm, primitive.hpp:58: pow(pow(x,2)+pow(y,2)+pow(z,2),0.5).
 * pow(x,2) is x*x (both correctly rounded); pow(s,0.5) is written as the
 * correctly rounded sqrt, which is what glibc's pow returns for it. */
static inline double vnorm(double x, double y, double z)
{
    return sqrt((x * x + y * y) + z * z);
}

/* ramp, primitive.cpp:122-129 */
static inline double ramp(double t, double t_threshold)
{
    if (t < 0) return 0.0;
    else if (t >= 0 && t < t_threshold) return t / t_threshold;
    return 1.0;
}

static inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

double oracle_jitter(uint32_t i, uint32_t j, int a, double r_p)
{
    uint64_t h = splitmix64(((uint64_t)i << 32) | (uint64_t)j);
    h = splitmix64(h + (uint64_t)(a + 1));
    /* same form as cluster.cpp:238-240: double(rand()%1000)/1e3*2.0*r_p-r_p */
    return (double)(h % 1000u) / 1e3 * 2.0 * r_p - r_p;
}

/* ======================================================================
 * Cell grid and particle routines (src/cluster.cpp)
 * ====================================================================== */

void oracle_cell_grid(const oracle_params *p, int *n3, double *d3)
{
    /* cluster.cpp:26-32 */
    const double L[3] = {p->lx, p->ly, p->lz};
    for (int a = 0; a < 3; a++) {
        n3[a] = (int)floor(L[a] / p->r_c + 2 * p->margin / p->r_c);
        d3[a] = fmod(L[a] + 2 * p->margin, p->r_c) / n3[a] + p->r_c;
    }
}

static inline int cell_coord(double x, double dcell, double margin)
{
    /* cluster.cpp:140-142 */
    return (int)floor(x / dcell + margin / dcell);
}

int oracle_cells_build(const oracle_params *p, int n, const double *pos,
                       int *cell_start, int *ids, int *it_mark)
{
    int nc[3];
    double dc[3];
    oracle_cell_grid(p, nc, dc);
    const int ncell = nc[0] * nc[1] * nc[2];
    int *cnt = (int *)calloc((size_t)ncell + 1, sizeof(int));
    int *cellof = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    double y_max = 0.0;
    int rc = 0;
    for (int it = 0; it < n; it++) {
        /* cluster.cpp:134-138 */
        if (pos[3 * it + 1] > y_max) { y_max = pos[3 * it + 1]; *it_mark = it; }
        int i = cell_coord(pos[3 * it + 0], dc[0], p->margin);
        int j = cell_coord(pos[3 * it + 1], dc[1], p->margin);
        int k = cell_coord(pos[3 * it + 2], dc[2], p->margin);
        if (i < 0 || i >= nc[0] || j < 0 || j >= nc[1] || k < 0 || k >= nc[2]) {
            rc = -1;  /* cluster.cpp:144-163 */
            cellof[it] = -1;
            continue;
        }
        cellof[it] = i + j * nc[0] + k * nc[0] * nc[1];
        cnt[cellof[it]]++;
    }
    cell_start[0] = 0;
    for (int c = 0; c < ncell; c++) cell_start[c + 1] = cell_start[c] + cnt[c];
    memset(cnt, 0, sizeof(int) * (size_t)ncell);
    /* append in ascending id order: list::append, linked_list.cpp:30-43 */
    for (int it = 0; it < n; it++) {
        int c = cellof[it];
        if (c < 0) continue;
        ids[cell_start[c] + cnt[c]++] = it;
    }
    free(cnt);
    free(cellof);
    return rc;
}

/* Minimum-image displacement d = p_it - p_id, cluster.cpp:223-242. */
static inline void pair_disp(const oracle_params *p, const int *nc, const double *dc,
                             const double *pos, int it, int id, double *d)
{
    d[0] = pos[3 * it + 0] - pos[3 * id + 0];
    d[1] = pos[3 * it + 1] - pos[3 * id + 1];
    d[2] = pos[3 * it + 2] - pos[3 * id + 2];
    const double L[3] = {p->lx, p->ly, p->lz};
    for (int a = 0; a < 3; a++) {
        if (d[a] > dc[a] * (nc[a] - 2)) d[a] -= L[a];
        else if (d[a] < -dc[a] * (nc[a] - 2)) d[a] += L[a];
    }
    if (vnorm(d[0], d[1], d[2]) == 0.0) {
        for (int a = 0; a < 3; a++) d[a] = oracle_jitter((uint32_t)it, (uint32_t)id, a, p->r_p);
    }
}

/* Walk the 27-cell neighbourhood exactly as cluster.cpp:193-273 does and
 * call back for each id > it. */
typedef void (*pair_cb)(void *ctx, int it, int id);

static void walk_pairs(const oracle_params *p, int n, const double *pos,
                       const int *cell_start, const int *ids,
                       const int *nc, const double *dc, void *ctx, pair_cb cb)
{
    for (int it = 0; it < n; it++) {
        int o = cell_coord(pos[3 * it + 0], dc[0], p->margin);
        int pp = cell_coord(pos[3 * it + 1], dc[1], p->margin);
        int q = cell_coord(pos[3 * it + 2], dc[2], p->margin);
        for (int k = q - 1; k <= q + 1; k++)
            for (int j = pp - 1; j <= pp + 1; j++)
                for (int i = o - 1; i <= o + 1; i++) {
                    int co = (i + nc[0]) % nc[0];
                    co += (j + nc[1]) % nc[1] * nc[0];
                    co += (k + nc[2]) % nc[2] * nc[0] * nc[1];
                    for (int s = cell_start[co]; s < cell_start[co + 1]; s++) {
                        int id = ids[s];
                        if (id > it) cb(ctx, it, id);
                    }
                }
    }
}

typedef struct {
    const oracle_params *p;
    const double *pos;
    const int *nc;
    const double *dc;
    int64_t count, max_pairs;
    int *pairs;
} pairs_ctx;

static void pairs_cb(void *vctx, int it, int id)
{
    pairs_ctx *c = (pairs_ctx *)vctx;
    double d[3];
    pair_disp(c->p, c->nc, c->dc, c->pos, it, id, d);
    if (vnorm(d[0], d[1], d[2]) < c->p->r_eq) {
        if (c->count < c->max_pairs) {
            c->pairs[2 * c->count] = it;
            c->pairs[2 * c->count + 1] = id;
        }
        c->count++;
    }
}

int64_t oracle_contact_pairs(const oracle_params *p, int n, const double *pos,
                             int64_t max_pairs, int *pairs)
{
    int nc[3];
    double dc[3];
    oracle_cell_grid(p, nc, dc);
    int ncell = nc[0] * nc[1] * nc[2];
    int *cs = (int *)malloc(sizeof(int) * (size_t)(ncell + 1));
    int *ids = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int mark = -1;
    if (oracle_cells_build(p, n, pos, cs, ids, &mark) != 0) { free(cs); free(ids); return -1; }
    pairs_ctx c = {p, pos, nc, dc, 0, max_pairs, pairs};
    walk_pairs(p, n, pos, cs, ids, nc, dc, &c, pairs_cb);
    free(cs);
    free(ids);
    return c.count;
}

typedef struct {
    const oracle_params *p;
    const double *pos;
    const int *nc;
    const double *dc;
    const double *f_pi;
    int it_mark;
    double rmp;
    double *f_pp;
    double tau;
} force_ctx;

static void force_cb(void *vctx, int it, int id)
{
    force_ctx *c = (force_ctx *)vctx;
    double d[3];
    pair_disp(c->p, c->nc, c->dc, c->pos, it, id, d);
    double nrm = vnorm(d[0], d[1], d[2]);
    if (nrm < c->p->r_eq) {
        /* cluster.cpp:248-250: (r_eq-|d|)*ramp*k_n*(d/|d|) */
        double s = (c->p->r_eq - nrm) * c->rmp * c->p->k_n;
        double fn[3];
        for (int a = 0; a < 3; a++) fn[a] = (d[a] / nrm) * s;
        for (int a = 0; a < 3; a++) {
            c->f_pp[3 * it + a] = c->f_pp[3 * it + a] + fn[a];
            c->f_pp[3 * id + a] = c->f_pp[3 * id + a] - fn[a];
        }
        /* cluster.cpp:256-259 */
        if (c->f_pi && c->it_mark >= 0) {
            const double *a = c->f_pi + 3 * it, *b = c->f_pi + 3 * c->it_mark;
            if (vnorm(a[0], a[1], a[2]) != 0.0 && vnorm(b[0], b[1], b[2]) != 0.0)
                c->tau += 0.5 * ((fn[0] * d[0] + fn[1] * d[1]) + fn[2] * d[2]);
        }
    }
}

int oracle_contact_forces(const oracle_params *p, int n, const double *pos,
                          const double *f_pi, int it_mark, int t,
                          double *f_pp, double *tau)
{
    int nc[3];
    double dc[3];
    oracle_cell_grid(p, nc, dc);
    int ncell = nc[0] * nc[1] * nc[2];
    int *cs = (int *)malloc(sizeof(int) * (size_t)(ncell + 1));
    int *ids = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int mark_unused = -1;
    int rc = oracle_cells_build(p, n, pos, cs, ids, &mark_unused);
    for (int i = 0; i < 3 * n; i++) f_pp[i] = 0.0;  /* cluster.cpp:191 */
    force_ctx c = {p, pos, nc, dc, f_pi, it_mark, ramp((double)t, p->ramp_t), f_pp, 0.0};
    if (rc == 0) walk_pairs(p, n, pos, cs, ids, nc, dc, &c, force_cb);
    if (tau) *tau = c.tau;
    free(cs);
    free(ids);
    return rc;
}

static inline double dx_of(const oracle_params *p, int a)
{
    /* head.hpp:37-39 */
    return a == 0 ? p->lx / p->nx : (a == 1 ? p->ly / p->ny : p->lz / p->nz);
}

/* field::interpolate, field.cpp:223-248 (grid holds real values only) */
static double interp_one(const oracle_params *p, const double *v, size_t vstride,
                         double px, double py, double pz)
{
    const int Nx = p->nx, Ny = p->ny, Nz = p->nz;
    const double dx = dx_of(p, 0), dy = dx_of(p, 1), dz = dx_of(p, 2);
    int i = (int)floor(px / dx);
    int j = (int)floor(py / dy);
    int k = (int)floor(pz / dz);
    double m = px / dx - floor(px / dx);
    double n = py / dy - floor(py / dy);
    double o = pz / dz - floor(pz / dz);
#define V(ii, jj, kk) v[vstride * ((size_t)(ii) + (size_t)(jj) * Nx + (size_t)(kk) * Nx * Ny)]
    double r;
    r  = (1.0 - m) * (1.0 - n) * (1.0 - o) * V(i, j, k);
    r += (1.0 - m) * n * (1.0 - o) * V(i, (j + 1) % Ny, k);
    r += m * (1.0 - n) * (1.0 - o) * V((i + 1) % Nx, j, k);
    r += m * n * (1.0 - o) * V((i + 1) % Nx, (j + 1) % Ny, k);
    r += (1.0 - m) * (1.0 - n) * o * V(i, j, (k + 1) % Nz);
    r += (1.0 - m) * n * o * V(i, (j + 1) % Ny, (k + 1) % Nz);
    r += m * (1.0 - n) * o * V((i + 1) % Nx, j, (k + 1) % Nz);
    r += m * n * o * V((i + 1) % Nx, (j + 1) % Ny, (k + 1) % Nz);
#undef V
    return r;
}

void oracle_interp(const oracle_params *p, const double *grid, int n,
                   const double *pos, double *out)
{
    for (int it = 0; it < n; it++)
        out[it] = interp_one(p, grid, 1, pos[3 * it], pos[3 * it + 1], pos[3 * it + 2]);
}

/* field::interpolate2, field.cpp:251-281, size = 3 */
static double interp2_one(const oracle_params *p, const double *v, size_t vstride,
                          double px, double py, double pz)
{
    const int Nx = p->nx, Ny = p->ny, Nz = p->nz;
    const int size = 3;
    const double dx = dx_of(p, 0), dy = dx_of(p, 1), dz = dx_of(p, 2);
    int i = (int)floor(px / dx);
    int j = (int)floor(py / dy);
    int k = (int)floor(pz / dz);
    double m = px / dx - floor(px / dx);
    double n = py / dy - floor(py / dy);
    double o = pz / dz - floor(pz / dz);
    m = m / size + 1.0 / size;
    n = n / size + 1.0 / size;
    o = o / size + 1.0 / size;
    /* the reference's node indices: i-(1-size)/2 = i+1 (C integer division)
     * and i+(1+size)/2 = i+2, wrapped */
    const int i0 = (i - (1 - size) / 2 + Nx) % Nx, i1 = (i + (1 + size) / 2) % Nx;
    const int j0 = (j - (1 - size) / 2 + Ny) % Ny, j1 = (j + (1 + size) / 2) % Ny;
    const int k0 = (k - (1 - size) / 2 + Nz) % Nz, k1 = (k + (1 + size) / 2) % Nz;
#define V(ii, jj, kk) v[vstride * ((size_t)(ii) + (size_t)(jj) * Nx + (size_t)(kk) * Nx * Ny)]
    double r;
    r  = (1.0 - m) * (1.0 - n) * (1.0 - o) * V(i0, j0, k0);
    r += (1.0 - m) * n * (1.0 - o) * V(i0, j1, k0);
    r += m * (1.0 - n) * (1.0 - o) * V(i1, j0, k0);
    r += m * n * (1.0 - o) * V(i1, j1, k0);
    r += (1.0 - m) * (1.0 - n) * o * V(i0, j0, k1);
    r += (1.0 - m) * n * o * V(i0, j1, k1);
    r += m * (1.0 - n) * o * V(i1, j0, k1);
    r += m * n * o * V(i1, j1, k1);
#undef V
    return r;
}

void oracle_interp2(const oracle_params *p, const double *grid, int n,
                    const double *pos, double *out)
{
    for (int it = 0; it < n; it++)
        out[it] = interp2_one(p, grid, 1, pos[3 * it], pos[3 * it + 1], pos[3 * it + 2]);
}

/* element::get_f_rand, element.cpp:109-144, one component: the do-while
 * redraws while |g| > 5.0; rand() % 1000 becomes the counter-based draw of
 * the header */
#define NOISE_TRIES 64
static double brownian_g(uint64_t seed, int t, int stage, int it, int d, double pi)
{
    uint64_t key = splitmix64(seed);
    key = splitmix64(key ^ (uint64_t)(uint32_t)t);
    key = splitmix64(key ^ (((uint64_t)(uint32_t)(3 * (stage - 1) + d) << 32) | (uint64_t)(uint32_t)it));
    for (int k = 0; k < NOISE_TRIES; k++) {
        const uint64_t u = splitmix64(key + (uint64_t)(k + 1) * 0xD1B54A32D192ED03ull);
        const double r1 = (double)((u >> 32) % 1000u) / 1000;
        const double r2 = (double)((u & 0xffffffffull) % 1000u) / 1000;
        const double g = sqrt(-2 * log(r1)) * cos(2 * pi * r2);
        if (!(fabs(g) > 5.0)) return g;
    }
    return 0.0;
}

static inline double brownian_amp(const oracle_params *p)
{
    /* element.cpp:139-143 */
    return sqrt(12.0 * p->pi * p->r_p * p->mu * p->kt / p->dt);
}

void oracle_brownian(const oracle_params *p, int t, int stage, int n, double *out)
{
    const double amp = brownian_amp(p);
    for (int it = 0; it < n; it++)
        for (int d = 0; d < 3; d++)
            out[3 * it + d] = amp * brownian_g(p->noise_seed, t, stage, it, d, p->pi);
}

void oracle_drag(const oracle_params *p, int n, const double *u, const double *uf,
                 const double *radius, double *out)
{
    for (int it = 0; it < n; it++) {
        const double r = radius ? radius[it] : p->r_p;
        const double c = -6 * p->pi * r * p->mu;
        for (int d = 0; d < 3; d++) out[3 * it + d] = c * (u[3 * it + d] - uf[3 * it + d]);
    }
}

/* cluster::compute_f_r, cluster.cpp:344-410; out has stride `ostride`
 * doubles per grid point and component offset `ocomp`. */
static void spread_into(const oracle_params *p, int n, const double *pos,
                        const double *force, double *out, size_t ostride, size_t ocomp)
{
    const int Nx = p->nx, Ny = p->ny, Nz = p->nz;
    const double dx = dx_of(p, 0), dy = dx_of(p, 1), dz = dx_of(p, 2);
    for (int it = 0; it < n; it++) {
        double f[3];
        for (int a = 0; a < 3; a++) f[a] = force[3 * it + a] / dx / dy / dz;
        const double px = pos[3 * it], py = pos[3 * it + 1], pz = pos[3 * it + 2];
        int x = (int)floor(px / dx);
        int y = (int)floor(py / dy);
        int z = (int)floor(pz / dz);
        double o = px / dx - floor(px / dx);
        double pp = py / dy - floor(py / dy);
        double q = pz / dz - floor(pz / dz);
        int xs[8] = {x, (x + 1) % Nx, (x + 1) % Nx, (x + 1) % Nx, (x + 1) % Nx, x, x, x};
        int ys[8] = {y, y, (y + 1) % Ny, y, (y + 1) % Ny, y, (y + 1) % Ny, (y + 1) % Ny};
        int zs[8] = {z, z, z, (z + 1) % Nz, (z + 1) % Nz, (z + 1) % Nz, z, (z + 1) % Nz};
        for (int c = 0; c < 8; c++) {
            size_t co = (size_t)xs[c] + (size_t)ys[c] * Nx + (size_t)zs[c] * Nx * Ny;
            for (int a = 0; a < 3; a++) {
                double w;
                /* cluster.cpp:369-406, weights multiplied onto f left to right */
                switch (c) {
                case 0: w = f[a] * (1 - o) * (1 - pp) * (1 - q); break;
                case 1: w = f[a] * o * (1 - pp) * (1 - q); break;
                case 2: w = f[a] * o * pp * (1 - q); break;
                case 3: w = f[a] * o * (1 - pp) * q; break;
                case 4: w = f[a] * o * pp * q; break;
                case 5: w = f[a] * (1 - o) * (1 - pp) * q; break;
                case 6: w = f[a] * (1 - o) * pp * (1 - q); break;
                default: w = f[a] * (1 - o) * pp * q; break;
                }
                out[(size_t)a * ocomp + co * ostride] += w;   /* field::add_value, field.cpp:150-155 */
            }
        }
    }
}

void oracle_spread(const oracle_params *p, int n, const double *pos,
                   const double *force, double *out)
{
    size_t G = (size_t)p->nx * p->ny * p->nz;
    memset(out, 0, sizeof(double) * 3 * G);
    spread_into(p, n, pos, force, out, 1, G);
}

/* element::get_f_pi, element.cpp:75-97 */
static void f_pi_one(const oracle_params *p, double phi, const double *g,
                     double r, double cth, double *out)
{
    double f_mag = 0.0, d = 0.0;
    if (fabs(phi) <= p->tanh2) {
        double gn = vnorm(g[0], g[1], g[2]);
        if (gn != 0.0) {
            d = 0.5 * log((1.0 + phi) / (1.0 - phi)) * sqrt(2.0) * p->epsilon;
            if (cth != 0.0) d = d - r * cth;  /* wetting extension (off at 0) */
            if (fabs(d) < r) f_mag = -p->a_pi * p->pi * d * p->sigma;
            else f_mag = 0.0;
            for (int a = 0; a < 3; a++) out[a] = g[a] / gn * f_mag;
            return;
        }
    }
    out[0] = out[1] = out[2] = 0.0;
}

void oracle_f_pi(const oracle_params *p, int n, const double *c_loc,
                 const double *dcdx, const double *radius,
                 const double *cos_theta, double *f_pi)
{
    for (int it = 0; it < n; it++)
        f_pi_one(p, c_loc[it], dcdx + 3 * it, radius ? radius[it] : p->r_p,
                 cos_theta ? cos_theta[it] : 0.0, f_pi + 3 * it);
}

/* ======================================================================
 * Full time step, src/main.cpp:75-222
 * ====================================================================== */
struct oracle_sim {
    oracle_params p;
    size_t G;
    int t;
    int nt;
    /* complex fields, interleaved, 2G doubles per component */
    double *c, *c_new, *c_old, *f, *f_old, *conv, *conv_old, *pch;
    double *grad[3], *u[3], *ft[3];
    double *kx, *ky, *kz; /* wave numbers per index, field.cpp:108-115 */
    /* particles */
    int n;
    double *pos, *pos_old, *pu, *pu_old, *f_pp, *f_pi, *c_loc, *dcdx, *uf, *f_rand, *rad, *cth;
    int *cell_start, *ids;
    int it_mark;
    double tau;
    double *W; /* wetting source (calibration model) or NULL */
};

static double *cfield(size_t G)
{
    return (double *)calloc(2 * G, sizeof(double));
}

oracle_sim *oracle_create(const oracle_params *p, const double *phi0,
                          const double *pos, const double *radius,
                          const double *cos_theta)
{
    oracle_sim *s = (oracle_sim *)calloc(1, sizeof(oracle_sim));
    s->p = *p;
    s->G = (size_t)p->nx * p->ny * p->nz;
    s->nt = p->n_threads > 0 ? p->n_threads : omp_get_max_threads();
    size_t G = s->G;
    s->c = cfield(G); s->c_new = cfield(G); s->c_old = cfield(G);
    s->f = cfield(G); s->f_old = cfield(G);
    s->conv = cfield(G); s->conv_old = cfield(G); s->pch = cfield(G);
    for (int d = 0; d < 3; d++) { s->grad[d] = cfield(G); s->u[d] = cfield(G); s->ft[d] = cfield(G); }
    for (size_t it = 0; it < G; it++) s->c[2 * it] = phi0[it];
    s->kx = (double *)malloc(sizeof(double) * p->nx);
    s->ky = (double *)malloc(sizeof(double) * p->ny);
    s->kz = (double *)malloc(sizeof(double) * p->nz);
    /* field.cpp:108-115: i*2.0*pi/Lx, with i <= N/2 kept positive */
    for (int i = 0; i < p->nx; i++) s->kx[i] = (i <= p->nx / 2 ? i : i - p->nx) * 2.0 * p->pi / p->lx;
    for (int j = 0; j < p->ny; j++) s->ky[j] = (j <= p->ny / 2 ? j : j - p->ny) * 2.0 * p->pi / p->ly;
    for (int k = 0; k < p->nz; k++) s->kz[k] = (k <= p->nz / 2 ? k : k - p->nz) * 2.0 * p->pi / p->lz;

    int n = p->switch_particle ? p->n_particles : 0;
    s->n = n;
    size_t n3 = 3 * (size_t)(n > 0 ? n : 1);
    s->pos = (double *)calloc(n3, sizeof(double));
    s->pos_old = (double *)calloc(n3, sizeof(double));
    s->pu = (double *)calloc(n3, sizeof(double));
    s->pu_old = (double *)calloc(n3, sizeof(double));
    s->f_pp = (double *)calloc(n3, sizeof(double));
    s->f_pi = (double *)calloc(n3, sizeof(double));
    s->dcdx = (double *)calloc(n3, sizeof(double));
    s->uf = (double *)calloc(n3, sizeof(double));
    s->f_rand = (double *)calloc(n3, sizeof(double));
    s->c_loc = (double *)calloc(n3 / 3, sizeof(double));
    s->rad = (double *)calloc(n3 / 3, sizeof(double));
    s->cth = (double *)calloc(n3 / 3, sizeof(double));
    for (int i = 0; i < n; i++) {
        s->c_loc[i] = 1.0;  /* element.cpp:34 */
        s->rad[i] = radius ? radius[i] : p->r_p;
        s->cth[i] = cos_theta ? cos_theta[i] : 0.0;
    }
    if (n > 0) {
        memcpy(s->pos, pos, sizeof(double) * 3 * n);
        memcpy(s->pos_old, pos, sizeof(double) * 3 * n);  /* cluster.cpp:88 */
    }
    int nc[3];
    double dc[3];
    oracle_cell_grid(p, nc, dc);
    s->cell_start = (int *)calloc((size_t)nc[0] * nc[1] * nc[2] + 1, sizeof(int));
    s->ids = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    s->it_mark = -1;  /* cluster.cpp:39 */
    /* calibration model: W = -a_w sum_p cos(theta_p) S_p, the first
     * component of the spread of (w_p, 0, 0) (cluster.cpp:344-410) */
    if (p->wetting != 0.0 && n > 0 && cos_theta) {
        double *w = (double *)calloc(3 * (size_t)n, sizeof(double));
        double *o = (double *)calloc(3 * G, sizeof(double));
        for (int i = 0; i < n; i++) w[3 * i] = -p->wetting * cos_theta[i];
        oracle_spread(p, n, s->pos, w, o);
        s->W = (double *)malloc(sizeof(double) * G);
        memcpy(s->W, o, sizeof(double) * G);
        free(w);
        free(o);
    }
    /* calibration model: the contact forces of the fixed particles at the
     * full ramp (t beyond ramp_t, primitive.cpp:122-129) */
    if (p->fix_particles && n > 0) {
        double tau = 0.0;
        const int tfull = (int)ceil(p->ramp_t) + 1;
        oracle_contact_forces(p, n, s->pos, NULL, -1, tfull, s->f_pp, &tau);
    }
    return s;
}

void oracle_destroy(oracle_sim *s)
{
    if (!s) return;
    free(s->c); free(s->c_new); free(s->c_old); free(s->f); free(s->f_old);
    free(s->conv); free(s->conv_old); free(s->pch);
    for (int d = 0; d < 3; d++) { free(s->grad[d]); free(s->u[d]); free(s->ft[d]); }
    free(s->kx); free(s->ky); free(s->kz);
    free(s->pos); free(s->pos_old); free(s->pu); free(s->pu_old); free(s->f_pp);
    free(s->f_pi); free(s->dcdx); free(s->uf); free(s->f_rand); free(s->c_loc); free(s->rad); free(s->cth);
    free(s->cell_start); free(s->ids);
    free(s->W);
    free(s);
}

static void fftc(oracle_sim *s, double *v, int sign)
{
    oracle_fft3d(s->p.nx, s->p.ny, s->p.nz, v, sign, s->nt);
}

/* field::ifft = backward FFT + normalize (field.cpp:447-455,
 * primitive.cpp:54-62: divide re and im by Nx*Ny*Nz) */
static void ifftc(oracle_sim *s, double *v)
{
    fftc(s, v, +1);
    const double Gd = (double)(s->p.nx * s->p.ny * s->p.nz);
    const size_t G = s->G;
    #pragma omp parallel for num_threads(s->nt) schedule(static)
    for (size_t it = 0; it < G; it++) {
        v[2 * it] /= Gd;
        v[2 * it + 1] /= Gd;
    }
}

/* index -> (i,j,k), field.cpp:106-108 */
#define IJK(it, i, j, k)                                   \
    const int k = (int)((it) / ((size_t)nx * ny));         \
    const int j = (int)((it) % ((size_t)nx * ny) / nx);    \
    const int i = (int)((it) % ((size_t)nx * ny) % nx)

/* f = (lambda/sqr(epsilon))*(c.power(3.0)-c), main.cpp:81,153 */
static void compute_f(oracle_sim *s)
{
    const double sc = s->p.lambda / (s->p.epsilon * s->p.epsilon);
    const size_t G = s->G;
    double *c = s->c, *f = s->f;
    #pragma omp parallel for num_threads(s->nt) schedule(static)
    for (size_t it = 0; it < G; it++) {
        f[2 * it] = (pow(c[2 * it], 3.0) - c[2 * it]) * sc;
        if (s->W) f[2 * it] += s->W[it];
        f[2 * it + 1] = 0.0;
    }
}

/* grad_c = c.grad(kf), field.cpp:288-328: i*k_d*c_hat, zero on the
 * Nyquist plane of axis d */
static void compute_grad(oracle_sim *s)
{
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    const size_t G = s->G;
    const double *c = s->c;
    #pragma omp parallel for num_threads(s->nt) schedule(static)
    for (size_t it = 0; it < G; it++) {
        IJK(it, i, j, k);
        const double kv[3] = {s->kx[i], s->ky[j], s->kz[k]};
        const int nyq[3] = {i == nx / 2, j == ny / 2, k == nz / 2};
        for (int d = 0; d < 3; d++) {
            if (nyq[d]) {
                s->grad[d][2 * it] = 0.0;
                s->grad[d][2 * it + 1] = 0.0;
            } else {
                s->grad[d][2 * it] = -kv[d] * c[2 * it + 1];
                s->grad[d][2 * it + 1] = kv[d] * c[2 * it];
            }
        }
    }
}

static inline double k2_at(const oracle_sim *s, int i, int j, int k)
{
    /* kf.power(2.0), field.cpp:421-428: sum of pow(k_d,2) from 0.0 */
    double a = 0.0;
    a += s->kx[i] * s->kx[i];
    a += s->ky[j] * s->ky[j];
    a += s->kz[k] * s->kz[k];
    return a;
}

static void particle_stage(oracle_sim *s, int stage, int tp1, int *err);

/* f_total.fft; u = f_total - kf*(kf.dot(f_total))/kf2; u /= mu*kf2;
 * u.ifft  (main.cpp:175-179; dot field.cpp:375-389, division by zero gives
 * zero field.cpp:1010-1019, 759-768) */
static void stokes_solve(oracle_sim *s)
{
    const oracle_params *P = &s->p;
    const int nx = P->nx, ny = P->ny;
    const size_t G = s->G;
    for (int d = 0; d < 3; d++) fftc(s, s->ft[d], -1);
    #pragma omp parallel for num_threads(s->nt) schedule(static)
    for (size_t it = 0; it < G; it++) {
        IJK(it, i, j, k);
        const double kv[3] = {s->kx[i], s->ky[j], s->kz[k]};
        const double k2 = k2_at(s, i, j, k);
        double dr = 0.0, di = 0.0;
        for (int d = 0; d < 3; d++) {
            dr += kv[d] * s->ft[d][2 * it];
            di += kv[d] * s->ft[d][2 * it + 1];
        }
        const double vis = k2 * P->mu;
        for (int d = 0; d < 3; d++) {
            double pr = kv[d] * dr, pi_ = kv[d] * di;
            if (k2 != 0.0) { pr /= k2; pi_ /= k2; } else { pr = 0.0; pi_ = 0.0; }
            double ur = s->ft[d][2 * it] - pr, ui = s->ft[d][2 * it + 1] - pi_;
            if (vis != 0.0) { ur /= vis; ui /= vis; } else { ur = 0.0; ui = 0.0; }
            s->u[d][2 * it] = ur;
            s->u[d][2 * it + 1] = ui;
        }
    }
    for (int d = 0; d < 3; d++) ifftc(s, s->u[d]);
}

void oracle_stokes(const oracle_params *p, const double *F, double *u)
{
    oracle_params q = *p;
    q.switch_particle = 0;
    q.n_particles = 0;
    double *zero = (double *)calloc((size_t)p->nx * p->ny * p->nz, sizeof(double));
    oracle_sim *s = oracle_create(&q, zero, NULL, NULL, NULL);
    for (int d = 0; d < 3; d++)
        for (size_t it = 0; it < s->G; it++) {
            s->ft[d][2 * it] = F[(size_t)d * s->G + it];
            s->ft[d][2 * it + 1] = 0.0;
        }
    stokes_solve(s);
    oracle_get_u(s, u);
    oracle_destroy(s);
    free(zero);
}

int oracle_step(oracle_sim *s, int nsteps)
{
    const oracle_params *P = &s->p;
    const int nx = P->nx, ny = P->ny;
    const size_t G = s->G;
    int err = 0;
    for (int step = 0; step < nsteps; step++) {
        const int t = s->t;
        /* ---------------- Phase field, main.cpp:79-141 ---------------- */
        if (P->switch_ch) {
            compute_f(s);                        /* :81 */
            fftc(s, s->f, -1);                   /* :83 */
            fftc(s, s->c, -1);                   /* :84 */
            if (P->switch_ns) {
                compute_grad(s);                 /* :88 */
                for (int d = 0; d < 3; d++) ifftc(s, s->grad[d]);  /* :89 */
                /* :91 conv_c = u.dot(grad_c), real space (field.cpp:392-400) */
                #pragma omp parallel for num_threads(s->nt) schedule(static)
                for (size_t it = 0; it < G; it++) {
                    double a = 0.0;
                    for (int d = 0; d < 3; d++) a += s->u[d][2 * it] * s->grad[d][2 * it];
                    s->conv[2 * it] = a;
                    s->conv[2 * it + 1] = 0.0;
                }
                fftc(s, s->conv, -1);            /* :92 */
            }
            if (t == 0) {
                /* :98-113 first-order IMEX start */
                memcpy(s->c_old, s->c, sizeof(double) * 2 * G);
                memcpy(s->f_old, s->f, sizeof(double) * 2 * G);
                if (P->switch_ns) memcpy(s->conv_old, s->conv, sizeof(double) * 2 * G);
                const double a1 = P->dt * P->mobility;
                const double a2 = P->dt * P->mobility * P->lambda;
                #pragma omp parallel for num_threads(s->nt) schedule(static)
                for (size_t it = 0; it < G; it++) {
                    IJK(it, i, j, k);
                    const double k2 = k2_at(s, i, j, k);
                    const double m = k2 * a1;
                    double cr = s->c[2 * it], ci = s->c[2 * it + 1];
                    cr -= m * s->f[2 * it];
                    ci -= m * s->f[2 * it + 1];
                    if (P->switch_ns) {
                        cr -= s->conv[2 * it] * P->dt;
                        ci -= s->conv[2 * it + 1] * P->dt;
                    }
                    const double div = k2 * k2 * a2 + 1.0;
                    if (div != 0.0) { cr /= div; ci /= div; } else { cr = 0.0; ci = 0.0; }
                    s->c[2 * it] = cr;
                    s->c[2 * it + 1] = ci;
                }
            } else {
                /* :117-137 SBDF2 */
                const double a4 = 4.0 * P->dt * P->mobility;
                const double a2 = 2.0 * P->dt * P->mobility;
                const double b4 = 4.0 * P->dt;
                const double b2 = 2.0 * P->dt;
                const double dd = 2.0 * P->dt * P->mobility * P->lambda;
                #pragma omp parallel for num_threads(s->nt) schedule(static)
                for (size_t it = 0; it < G; it++) {
                    IJK(it, i, j, k);
                    const double k2 = k2_at(s, i, j, k);
                    double nr = s->c[2 * it] * 4.0 - s->c_old[2 * it];
                    double ni = s->c[2 * it + 1] * 4.0 - s->c_old[2 * it + 1];
                    const double m4 = k2 * a4, m2 = k2 * a2;
                    nr -= m4 * s->f[2 * it];
                    ni -= m4 * s->f[2 * it + 1];
                    nr += m2 * s->f_old[2 * it];
                    ni += m2 * s->f_old[2 * it + 1];
                    if (P->switch_ns) {
                        nr -= s->conv[2 * it] * b4;
                        ni -= s->conv[2 * it + 1] * b4;
                        nr += s->conv_old[2 * it] * b2;
                        ni += s->conv_old[2 * it + 1] * b2;
                    }
                    const double div = k2 * k2 * dd + 3.0;
                    if (div != 0.0) { nr /= div; ni /= div; } else { nr = 0.0; ni = 0.0; }
                    s->c_new[2 * it] = nr;
                    s->c_new[2 * it + 1] = ni;
                }
                /* c_old = c; f_old = f; conv_c_old = conv_c; c = c_new */
                double *tmp = s->c_old; s->c_old = s->c; s->c = s->c_new; s->c_new = tmp;
                tmp = s->f_old; s->f_old = s->f; s->f = tmp;
                if (P->switch_ns) { tmp = s->conv_old; s->conv_old = s->conv; s->conv = tmp; }
            }
            ifftc(s, s->c);                      /* :140 */
        }

        /* ---------------- Fluid velocity, main.cpp:146-181 ------------- */
        if (P->switch_ns) {
            for (int d = 0; d < 3; d++) memset(s->ft[d], 0, sizeof(double) * 2 * G);  /* :149 */
            if (P->switch_ch) {
                compute_f(s);                    /* :153 */
                fftc(s, s->f, -1);               /* :154 */
                fftc(s, s->c, -1);               /* :155 */
                /* :157 pch = lambda*kf2*c + f */
                #pragma omp parallel for num_threads(s->nt) schedule(static)
                for (size_t it = 0; it < G; it++) {
                    IJK(it, i, j, k);
                    const double m = k2_at(s, i, j, k) * P->lambda;
                    s->pch[2 * it] = m * s->c[2 * it] + s->f[2 * it];
                    s->pch[2 * it + 1] = m * s->c[2 * it + 1] + s->f[2 * it + 1];
                }
                ifftc(s, s->pch);                /* :158 */
                compute_grad(s);                 /* :160 */
                for (int d = 0; d < 3; d++) ifftc(s, s->grad[d]);  /* :161 */
                ifftc(s, s->c);                  /* :162 */
                /* :164 f_total += pch*grad_c */
                #pragma omp parallel for num_threads(s->nt) schedule(static)
                for (size_t it = 0; it < G; it++)
                    for (int d = 0; d < 3; d++) {
                        s->ft[d][2 * it] = s->ft[d][2 * it] + s->pch[2 * it] * s->grad[d][2 * it];
                        s->ft[d][2 * it + 1] = 0.0;
                    }
            }
            if (P->switch_particle && s->n > 0) {
                /* :169 f_total += compute_f_r(), with f = f_pp + f_ext (f_ext = 0) */
                double *fr = (double *)calloc(3 * G, sizeof(double));
                double *fsum = (double *)malloc(sizeof(double) * 3 * s->n);
                for (int i = 0; i < 3 * s->n; i++) fsum[i] = s->f_pp[i] + 0.0;
                spread_into(P, s->n, s->pos, fsum, fr, 1, G);
                #pragma omp parallel for num_threads(s->nt) schedule(static)
                for (size_t it = 0; it < G; it++)
                    for (int d = 0; d < 3; d++) s->ft[d][2 * it] += fr[(size_t)d * G + it];
                free(fr);
                free(fsum);
            }
            /* :172 f_total += ((1.0+c)/2.0*rho1 + (1.0-c)/2.0*rho2)*g, g = (0,0,gravity) */
            {
                const double g3[3] = {0.0, 0.0, P->gravity};
                #pragma omp parallel for num_threads(s->nt) schedule(static)
                for (size_t it = 0; it < G; it++) {
                    const double cr = s->c[2 * it];
                    const double b = (cr + 1.0) / 2.0 * P->rho1 + (cr * (-1.0) + 1.0) / 2.0 * P->rho2;
                    for (int d = 0; d < 3; d++) {
                        s->ft[d][2 * it] += b * g3[d];
                        s->ft[d][2 * it + 1] = 0.0;
                    }
                }
            }
            stokes_solve(s);                     /* :175-179 */
        }

        /* ---------------- Particles, main.cpp:186-219 ------------------ */
        if (P->switch_particle && s->n > 0 && !P->fix_particles) {
            if (P->switch_ch) {
                fftc(s, s->c, -1);               /* :190 */
                compute_grad(s);                 /* :191 */
                for (int d = 0; d < 3; d++) ifftc(s, s->grad[d]);  /* :192 */
                ifftc(s, s->c);                  /* :193 */
            }
            particle_stage(s, 1, t + 1, &err);   /* :198-206 */
            particle_stage(s, 2, t + 1, &err);   /* :209-217 */
        }
        s->t++;
        if (err) return -1;
    }
    return 0;
}

/* One Heun stage: interpolate_local_variables, sorting, compute_f_pi,
 * compute_f_pp, compute_f_ext, advect, enforce_bc (main.cpp:198-217). */
static void particle_stage(oracle_sim *s, int stage, int tp1, int *err)
{
    const oracle_params *P = &s->p;
    const int n = s->n;
    /* cluster.cpp:307-323 -> element.cpp:38-71: uf = 0 (element.cpp:40),
     * or u interpolated (element.cpp:43-49 with interpolate or
     * interpolate2) when interp_u is set */
    for (int it = 0; it < n; it++) {
        const double *pp = s->pos + 3 * it;
        for (int d = 0; d < 3; d++) {
            double v = 0.0;
            if (P->switch_ns && P->interp_u == 1) v = interp_one(P, s->u[d], 2, pp[0], pp[1], pp[2]);
            else if (P->switch_ns && P->interp_u == 2) v = interp2_one(P, s->u[d], 2, pp[0], pp[1], pp[2]);
            s->uf[3 * it + d] = v;
        }
        if (P->switch_ch) {
            s->c_loc[it] = interp_one(P, s->c, 2, pp[0], pp[1], pp[2]);
            for (int d = 0; d < 3; d++)
                s->dcdx[3 * it + d] = interp_one(P, s->grad[d], 2, pp[0], pp[1], pp[2]);
        }
    }
    /* sorting, cluster.cpp:116-169 */
    if (oracle_cells_build(P, n, s->pos, s->cell_start, s->ids, &s->it_mark) != 0) *err = 1;
    /* compute_f_pi, cluster.cpp:286-290 */
    oracle_f_pi(P, n, s->c_loc, s->dcdx, s->rad, s->cth, s->f_pi);
    /* compute_f_pp, cluster.cpp:174-275 */
    {
        int nc[3];
        double dc[3];
        oracle_cell_grid(P, nc, dc);
        for (int i = 0; i < 3 * n; i++) s->f_pp[i] = 0.0;
        force_ctx c = {P, s->pos, nc, dc, s->f_pi, s->it_mark, ramp((double)tp1, P->ramp_t), s->f_pp, 0.0};
        if (!*err) walk_pairs(P, n, s->pos, s->cell_start, s->ids, nc, dc, &c, force_cb);
        s->tau = c.tau;
    }
    /* compute_f_ext: f_ext = 0 (element.cpp:100-104), or the Brownian
     * force when kt > 0 (element.cpp:109-144); advect, element.cpp:156-173 */
    if (P->kt > 0.0) oracle_brownian(P, tp1 - 1, stage, n, s->f_rand);
    for (int it = 0; it < n; it++) {
        const double r = s->rad[it];
        double v[3];
        for (int a = 0; a < 3; a++) {
            const double fext = P->kt > 0.0 ? s->f_rand[3 * it + a] : 0.0;
            double fsum = s->f_pi[3 * it + a] + fext + s->f_pp[3 * it + a];
            v[a] = s->uf[3 * it + a] + fsum / 6.0 / P->pi / P->mu / r;
        }
        if (stage == 1) {
            for (int a = 0; a < 3; a++) {
                s->pu_old[3 * it + a] = v[a];
                s->pos_old[3 * it + a] = s->pos[3 * it + a];
                s->pos[3 * it + a] = s->pos[3 * it + a] + v[a] * P->dt;
            }
        } else {
            const double hdt = 0.5 * P->dt;
            for (int a = 0; a < 3; a++) {
                s->pu[3 * it + a] = v[a];
                s->pos[3 * it + a] = s->pos_old[3 * it + a] + (s->pu_old[3 * it + a] + v[a]) * hdt;
            }
        }
        /* enforce_bc, element.cpp:177-186 */
        s->pos[3 * it + 0] = fmod(s->pos[3 * it + 0] + P->lx, P->lx);
        s->pos[3 * it + 1] = fmod(s->pos[3 * it + 1] + P->ly, P->ly);
        s->pos[3 * it + 2] = fmod(s->pos[3 * it + 2] + P->lz, P->lz);
    }
}

int oracle_time(const oracle_sim *s) { return s->t; }

static void get_real(const oracle_sim *s, double *const *src, int dim, double *out)
{
    for (int d = 0; d < dim; d++)
        for (size_t it = 0; it < s->G; it++) out[(size_t)d * s->G + it] = src[d][2 * it];
}

void oracle_get_phi(const oracle_sim *s, double *out) { double *v[1] = {s->c}; get_real(s, v, 1, out); }
void oracle_get_u(const oracle_sim *s, double *out) { get_real(s, s->u, 3, out); }
void oracle_get_grad_c(const oracle_sim *s, double *out) { get_real(s, s->grad, 3, out); }

void oracle_set_u(oracle_sim *s, const double *u)
{
    for (int d = 0; d < 3; d++)
        for (size_t it = 0; it < s->G; it++) {
            s->u[d][2 * it] = u[(size_t)d * s->G + it];
            s->u[d][2 * it + 1] = 0.0;
        }
}

void oracle_get_particles(const oracle_sim *s, double *p, double *p_old,
                          double *u, double *u_old, double *f_pp,
                          double *f_pi, double *c_loc)
{
    size_t b = sizeof(double) * 3 * (size_t)s->n;
    if (p) memcpy(p, s->pos, b);
    if (p_old) memcpy(p_old, s->pos_old, b);
    if (u) memcpy(u, s->pu, b);
    if (u_old) memcpy(u_old, s->pu_old, b);
    if (f_pp) memcpy(f_pp, s->f_pp, b);
    if (f_pi) memcpy(f_pi, s->f_pi, b);
    if (c_loc) memcpy(c_loc, s->c_loc, b / 3);
}

void oracle_get_particle_fluid(const oracle_sim *s, double *uf, double *f_rand)
{
    if (uf) memcpy(uf, s->uf, sizeof(double) * 3 * (size_t)s->n);
    if (f_rand) memcpy(f_rand, s->f_rand, sizeof(double) * 3 * (size_t)s->n);
}

double oracle_get_tau(const oracle_sim *s) { return s->tau; }
