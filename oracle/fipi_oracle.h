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
 *   Heun stages + wrap    two-body closed form with boundary crossings
 *   step order            convection from u^n (single mode, uniform u,
 *                         closed form); free energy decreases in pure CH
 *   calibration model     the wetting source and fixed particles in
 *                         closed form
 *   oracle_interp2        closed form on affine fields (the +1 node offset
 *                         of the reference), constants reproduced
 *   oracle_brownian       the exact distribution of g (all 10^6 pairs
 *                         r1, r2 enumerated), uniform draws
 *   oracle_drag, uf       uniform u: uf = U and f_drag = -(f_pi + f_pp)
 *   full step             independent NumPy restatement (tests/golden)
 * The reference itself needs FFTW, which this image lacks, so it cannot be
 * run ("parity unpinned" for FFTW's own round-off, which no check can
 * reach).
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
    /* The physics the reference carries but main() never calls (0, 0 give
     * the reference step):
     *   interp_u: the local fluid velocity uf of element::advect
     *   (element.cpp:162,168) -- 0: zero (element.cpp:40, the reference),
     *   1: field::interpolate of u (the commented code, element.cpp:43-49),
     *   2: field::interpolate2 of u (field.cpp:251-281, the "interpolation
     *   with wider support" that comment names);
     *   kt > 0: the Brownian force of element::get_f_rand
     *   (element.cpp:109-144) in the f_ext slot of the advection's force sum
     *   (f_pi + f_rand + f_pp), drawn per step, Heun stage, particle and
     *   component from noise_seed (oracle_brownian); it is not spread. */
    int    interp_u;
    double kt;
    unsigned long long noise_seed;
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
/* field::interpolate2, field.cpp:251-281 (kernel size 3), of a real grid
 * to n particles.  As the reference writes it, the two nodes of each axis
 * are floor(x/dx)+1 and floor(x/dx)+2 ((1-size)/2 = -1 is subtracted) with
 * weights 1-m', m', m' = (m+1)/3 for the fractional position m. */
void oracle_interp2(const oracle_params *p, const double *grid, int n,
                    const double *pos, double *out);
/* element::get_f_rand (element.cpp:109-144) of n particles at step t
 * (0-based) and Heun stage 1 or 2, out[3n]: sqrt(12 pi r_p mu kt/dt) g with
 * g = sqrt(-2 log r1) cos(2 pi r2), r1, r2 from {0, .., 999}/1000, drawn
 * again while |g| > 5.  rand() % 1000 is replaced by a counter-based draw:
 * attempt k (0-based, at most 64; g = 0 after that) of component d takes
 *   key = h(h(h(seed) ^ t) ^ ((3 (stage-1) + d) << 32 | it)),
 *   u = h(key + (k+1) * 0xD1B54A32D192ED03),
 *   r1 = (u >> 32) % 1000, r2 = (u mod 2^32) % 1000,
 * h the splitmix64 finaliser (x + 0x9E3779B97F4A7C15, then the two
 * xor-shift-multiply rounds and a final xor-shift). */
void oracle_brownian(const oracle_params *p, int t, int stage, int n, double *out);
/* element::get_f_drag (element.cpp:147-151): -6 pi r mu (u - uf) for n
 * particles, vectors [3n]; radius NULL: r_p. */
void oracle_drag(const oracle_params *p, int n, const double *u, const double *uf,
                 const double *radius, double *out);

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
/* uf and f_rand (3n each, either may be NULL) of the last Heun stage:
 * zero unless interp_u / kt are set. */
void oracle_get_particle_fluid(const oracle_sim *s, double *uf, double *f_rand);
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
