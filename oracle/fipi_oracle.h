/*
 * CPU FP64 oracle for the FIPI per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference time step (src/main.cpp:75-222 of the FIPI sources) and of the
 * routines it calls.  It is the checker the GPU library is compared with;
 * it is never linked into, loaded by or called from the product library
 * (fipi_amd/libfipi.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * Every function cites the reference file:line it restates.  Arithmetic
 * is written in the reference's operation order and compiled with
 * -ffp-contract=off so that each statement rounds as the reference's
 * x86-64 build does.
 *
 * Parity status of each part (how it is pinned):
 *   oracle_fft3d          brute-force DFT (4^3..8^3) and numpy.fft (pocketfft)
 *   wave numbers / grad   single Fourier modes -> analytic derivatives
 *   CH update             linear single-mode SBDF2 recurrence (closed form)
 *   Stokes projection     k.u_hat = 0, u_hat(0) = 0, single-mode closed form
 *   oracle_interp         exact on affine fields, weights sum to one
 *   oracle_spread         adjointness with oracle_interp, force conservation
 *   cell list / pairs     brute-force O(N^2) enumeration
 *   contact forces        two-body closed form, Newton's third law
 *   f_pi                  planar tanh profile gives d = x
 *   full step             independent NumPy restatement (tests/golden)
 * The composite step ordering is pinned only by the line-by-line
 * restatement and the independent NumPy restatement: the reference itself
 * needs FFTW, which this image lacks, so it cannot be run ("parity
 * unpinned" for FFTW's own round-off, which no check can reach).
 */
#ifndef FIPI_ORACLE_H
#define FIPI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical and numerical parameters.  Field meaning follows
 * src/head.hpp:27-74 and src/parameter.hpp:18-20; the reference holds them
 * as compile-time constants, here they are runtime values. */
typedef struct oracle_params {
    int    nx, ny, nz;        /* head.hpp:29-31 */
    int    n_particles;       /* head.hpp:63 */
    double lx, ly, lz;        /* head.hpp:33-35 */
    double pi;                /* head.hpp:27 (3.1415926 in the reference) */
    double dt;                /* head.hpp:41 */
    double ramp_t;            /* 0.2*t_end, cluster.cpp:249 */
    double epsilon;           /* head.hpp:46 */
    double lambda;            /* head.hpp:47 */
    double mobility;          /* head.hpp:58 */
    double mu;                /* head.hpp:54 (fluid viscosity) */
    double sigma;             /* head.hpp:45 */
    double a_pi;              /* head.hpp:57 (A) */
    double k_n;               /* parameter.hpp:18 */
    double r_p;               /* head.hpp:59 (particle radius) */
    double r_eq;              /* head.hpp:64 */
    double r_c;               /* head.hpp:65 (cell size target) */
    double margin;            /* parameter.hpp:20 */
    double rho1, rho2;        /* head.hpp:52-53 */
    double gravity;           /* head.hpp:51 */
    double tanh2;             /* tanh(2.0), element.cpp:84 */
    int    switch_particle;   /* head.hpp:72 */
    int    switch_ch;         /* head.hpp:73 */
    int    switch_ns;         /* head.hpp:74 */
    int    n_threads;         /* OpenMP threads used by the oracle */
    /* Calibration model of libfipi's adjoint (no counterpart in the
     * reference; 0 and 0 give the reference step):
     *   wetting a_w: f = (lambda/eps^2)(c^3 - c) + W with
     *   W = -a_w sum_p cos(theta_p) S_p, S_p the spread kernel of particle p
     *   over dx^3 (cos_theta of oracle_create);
     *   fix_particles: the particles stay where they are; their contact
     *   forces at the full ramp are computed once and spread every step. */
    double wetting;
    int    fix_particles;
} oracle_params;

/* ---- FFT (replaces FFTW's fftw_execute_dft, field.cpp:438-455) -------- */
/* In-place 3-D complex DFT of data[2*G] (interleaved re,im) with layout
 * it = i + nx*(j + ny*k).  sign = -1 forward, +1 backward, unnormalised
 * (FFTW's convention).  Sizes with prime factors 2, 3, 5 (and any other
 * prime through a generic odd-radix butterfly). */
int oracle_fft3d(int nx, int ny, int nz, double *data, int sign, int n_threads);
/* Factor order of the transforms made after the call (process-wide):
 * 0 = radix 4 then 2 (default), 1 = radix 2 only.  Both compute the DFT;
 * they differ in round-off only. */
void oracle_set_fft_order(int order);

/* ---- cell grid, cluster.cpp:21-40 ------------------------------------- */
void oracle_cell_grid(const oracle_params *p, int *ncell3, double *dcell3);

/* ---- particle kernels on caller arrays (positions in AoS x,y,z) -------- */
/* cluster::sorting, cluster.cpp:116-169: CSR cell list, ids ascending in
 * each cell.  cell_start has ncell+1 entries, ids has n entries.
 * Returns 0, or -1 if a particle is outside the sorting range
 * (the reference prints and exits there). it_mark gets argmax-y. */
int oracle_cells_build(const oracle_params *p, int n, const double *pos,
                       int *cell_start, int *ids, int *it_mark);

/* Contact pairs (i<j) found by cluster::compute_f_pp's 27-cell search,
 * cluster.cpp:193-262, in the reference's discovery order.  Writes up to
 * max_pairs pairs; returns the total number found. */
int64_t oracle_contact_pairs(const oracle_params *p, int n, const double *pos,
                             int64_t max_pairs, int *pairs);

/* cluster::compute_f_pp, cluster.cpp:174-275.  f_pi (3n) feeds the
 * packing-energy diagnostic tau (cluster.cpp:256-259). */
int oracle_contact_forces(const oracle_params *p, int n, const double *pos,
                          const double *f_pi, int it_mark, int t,
                          double *f_pp, double *tau);

/* field::interpolate, field.cpp:223-248, of one real grid field. */
void oracle_interp(const oracle_params *p, const double *grid, int n,
                   const double *pos, double *out);

/* cluster::compute_f_r, cluster.cpp:344-410: spread (f_pp+f_ext)/dx^3
 * of n particles to a 3-component real grid out[3*G] (component-major). */
void oracle_spread(const oracle_params *p, int n, const double *pos,
                   const double *force, double *out);

/* element::get_f_pi, element.cpp:75-97, for n particles.  radius may be
 * NULL (all r_p).  cos_theta may be NULL (neutral wetting, the reference);
 * otherwise the attachment spring is centred on d0 = r*cos_theta, which is
 * this project's wetting extension and reduces to the reference at 0. */
void oracle_f_pi(const oracle_params *p, int n, const double *c_loc,
                 const double *dcdx, const double *radius,
                 const double *cos_theta, double *f_pi);

/* Stokes block alone (main.cpp:175-179): u (3G, component-major) from a
 * real force field F (3G). */
void oracle_stokes(const oracle_params *p, const double *F, double *u);

/* ---- full time step (main.cpp:75-222) ---------------------------------- */
typedef struct oracle_sim oracle_sim;

/* phi0: G real values; pos: 3n; radius, cos_theta: n or NULL (all r_p,
 * neutral wetting). */
oracle_sim *oracle_create(const oracle_params *p, const double *phi0,
                          const double *pos, const double *radius,
                          const double *cos_theta);
void   oracle_destroy(oracle_sim *s);
/* Advance nsteps reference time steps. Returns 0 or -1 (sorting range). */
int    oracle_step(oracle_sim *s, int nsteps);
int    oracle_time(const oracle_sim *s);
/* Real parts of c (G), u (3G component-major), grad_c (3G). */
void   oracle_get_phi(const oracle_sim *s, double *out);
void   oracle_get_u(const oracle_sim *s, double *out);
void   oracle_get_grad_c(const oracle_sim *s, double *out);
/* Particle state, each 3n: p, p_old, u, u_old, f_pp, f_pi; c_loc (n). */
void   oracle_get_particles(const oracle_sim *s, double *p, double *p_old,
                            double *u, double *u_old, double *f_pp,
                            double *f_pi, double *c_loc);
double oracle_get_tau(const oracle_sim *s);
/* Replace u (3G real) -- lets tests start a step from an arbitrary flow. */
void   oracle_set_u(oracle_sim *s, const double *u);

/* Deterministic replacement for the rand() jitter of coincident particles
 * (cluster.cpp:234-242): component a (0..2) for the ordered pair i<j. */
double oracle_jitter(uint32_t i, uint32_t j, int a, double r_p);

#ifdef __cplusplus
}
#endif
#endif
