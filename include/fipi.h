/*
 * fipi.h -- C ABI of the MI355X-native FIPI time step (libfipi.so).
 *
 * The reference (FIPI, a serial C++ code) has no plugin or FFI layer: its
 * boundary is the C++ class API that src/main.cpp drives (`field`,
 * `cluster`, `element`, `list`) plus FFTW plan calls.  This header is the
 * thin C layer those classes call instead (the C++ facade in
 * fipi_amd/facade/ keeps the reference class names).  Plain pointers and
 * sizes only; every call returns a status code (0 = success) and never
 * exits, where the reference prints and calls exit(0) on misuse.
 *
 * Layouts
 *   real grid field   double[nz][ny][nx], x fastest (it = i + nx*(j + ny*k),
 *                     src/field.cpp:88-92)
 *   vector field      3 such arrays back to back (component-major)
 *   particle vectors  double[n][3] (x, y, z), as `vector` in primitive.hpp:37-43
 * Host pointers are copied in and out; "device" pointers are HIP device
 * memory used on the context's stream.
 */
#ifndef FIPI_H
#define FIPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FIPI_OK 0
#define FIPI_ERR_ARG (-1)         /* invalid argument or size */
#define FIPI_ERR_HIP (-2)         /* HIP runtime error */
#define FIPI_ERR_UNSUPPORTED (-3) /* grid size or option not built */
#define FIPI_ERR_RANGE (-4)       /* particle outside the sorting range (cluster.cpp:144-163) */
#define FIPI_ERR_STATE (-5)       /* call out of order (e.g. step before set_state) */
#define FIPI_ERR_OVERFLOW (-6)    /* a fixed-capacity buffer was too small */

#define FIPI_FP64 0        /* every field and particle quantity in FP64 */
#define FIPI_FP32_MIXED 1  /* FFTs, k-space update, interp/spread in FP32; k=0 mode,
                              mean phi, contact overlaps, positions in FP64 */

/* Runtime parameters.  The reference holds these as compile-time constants
 * (src/head.hpp:27-74, src/parameter.hpp:18-20); field names follow it. */
typedef struct fipi_params {
    int32_t nx, ny, nz;         /* grid, head.hpp:29-31 (each 16, 32, 48, 64, 80, 96, 128 or 256) */
    int32_t n_particles;        /* head.hpp:63 */
    double lx, ly, lz;          /* box, head.hpp:33-35 */
    double pi;                  /* head.hpp:27: the reference uses 3.1415926 */
    double dt;                  /* head.hpp:41 */
    double ramp_t;              /* contact ramp length 0.2*t_end, cluster.cpp:249 */
    double epsilon;             /* interface width, head.hpp:46 */
    double lambda;              /* gradient energy coefficient, head.hpp:47 */
    double mobility;            /* head.hpp:58 */
    double mu;                  /* fluid viscosity, head.hpp:54 */
    double sigma;               /* surface tension, head.hpp:45 */
    double a_pi;                /* attachment constant A, head.hpp:57 */
    double k_n;                 /* contact stiffness, parameter.hpp:18 */
    double r_p;                 /* particle radius, head.hpp:59 */
    double r_eq;                /* contact cutoff, head.hpp:64 */
    double r_c;                 /* contact cell size target, head.hpp:65 */
    double margin;              /* parameter.hpp:20 */
    double rho1, rho2;          /* head.hpp:52-53 */
    double gravity;             /* head.hpp:51 */
    double tanh2;               /* tanh(2.0), element.cpp:84 */
    int32_t switch_particle;    /* head.hpp:72 */
    int32_t switch_ch;          /* head.hpp:73 (must be 1) */
    int32_t switch_ns;          /* head.hpp:74 */
    int32_t precision;          /* FIPI_FP64 or FIPI_FP32_MIXED */
    /* Physics the reference carries but main() never calls; zero (a
     * zero-initialised struct) gives the reference step.  Single-box
     * contexts only (the ensemble and the adjoint refuse them). */
    int32_t interp_u;           /* local fluid velocity uf of element::advect (element.cpp:162,168):
                                   0 uf = 0 (element.cpp:40), 1 field::interpolate of u (the
                                   commented code, element.cpp:43-49), 2 field::interpolate2
                                   (field.cpp:251-281); needs switch_ns */
    double kt;                  /* thermal energy, head.hpp:68; kt > 0 adds the Brownian force of
                                   element::get_f_rand (element.cpp:109-144) in the f_ext slot of
                                   the advection's force sum (not spread onto the fluid) */
    uint64_t noise_seed;        /* seed of the Brownian draws (counter-based, replaces rand()) */
} fipi_params;

/* Diagnostics (printed by the reference at output steps). */
typedef struct fipi_diag {
    int32_t t;                  /* steps taken */
    int32_t it_mark;            /* argmax-y particle of the last sort, cluster.cpp:134-138 */
    double mean_phi;            /* FP64 mean of phi */
    double free_energy;         /* sum lambda(1/2|grad phi|^2 + 1/4 eps^-2 (phi^2-1)^2) dx^3,
                                   main.cpp:232 + field.cpp:506-516 */
    double interface_area;      /* sum |grad phi| dx^3 */
    double tau;                 /* packing energy, cluster.cpp:256-259,455 */
} fipi_diag;

typedef struct fipi_ctx fipi_ctx;

/* ---- context --------------------------------------------------------- */
/* Replaces the setup of main.cpp:29-73 (field/cluster construction and
 * FFTW plans).  Allocates all device state for one box on the current
 * HIP device and creates one stream. */
int fipi_create(const fipi_params *params, fipi_ctx **out);
int fipi_destroy(fipi_ctx *ctx);
/* Last error message of ctx (or of the calling thread when ctx is NULL). */
const char *fipi_last_error(const fipi_ctx *ctx);
/* Context stream (hipStream_t) as an opaque pointer. */
void *fipi_stream(fipi_ctx *ctx);

/* ---- state ----------------------------------------------------------- */
/* Initial state: phi (G reals, replaces field("SCALARS","phase_field"),
 * field.cpp:85-99), particle positions (3n, replaces cluster::distribute,
 * cluster.cpp:53-92), optional per-particle radius and cos(contact angle)
 * (NULL: r_p and neutral wetting, the reference).  Velocity is set to zero,
 * the step counter to 0. */
int fipi_set_state(fipi_ctx *ctx, const double *phi, const double *pos, const double *radius,
                   const double *cos_theta);
/* Replace the fluid velocity u (3G reals). */
int fipi_set_velocity(fipi_ctx *ctx, const double *u);
/* Copy out phi (G), u (3G), grad phi (3G). */
int fipi_get_phi(fipi_ctx *ctx, double *phi);
int fipi_get_velocity(fipi_ctx *ctx, double *u);
int fipi_get_grad_phi(fipi_ctx *ctx, double *grad);
/* Particle state, each array 3n (NULL to skip): position, previous
 * position, velocity, previous velocity, f_pp, f_pi (element.hpp:28-58). */
int fipi_get_particles(fipi_ctx *ctx, double *pos, double *pos_old, double *u, double *u_old, double *f_pp,
                       double *f_pi);
int fipi_get_time(fipi_ctx *ctx, int32_t *t);
int fipi_diagnostics(fipi_ctx *ctx, fipi_diag *out);

/* ---- the time step ---------------------------------------------------- */
/* Advance nsteps steps of the reference time loop body (main.cpp:75-222):
 * Cahn-Hilliard SBDF2 update, Stokes solve, two Heun particle stages.
 * Asynchronous on the context stream except for the error check at the
 * end of the call. */
int fipi_step(fipi_ctx *ctx, int32_t nsteps);
/* Same as fipi_step without the final synchronisation and error check
 * (for timing loops); errors are reported by the next fipi_sync. */
int fipi_step_async(fipi_ctx *ctx, int32_t nsteps);
int fipi_sync(fipi_ctx *ctx);

/* Per-kernel timing.  The reference's only instrumentation is clock()
 * around the whole run (main.cpp:25-26,248-249). */
typedef struct fipi_kernel_stat {
    char name[40];      /* launch site */
    int32_t launches;   /* launches timed */
    double total_ms;    /* summed HIP-event time on the context stream */
    double bytes;       /* algorithmic bytes (each array touched once), summed */
} fipi_kernel_stat;
/* Run nsteps steps (advancing the state like fipi_step) with HIP events
 * around every launch; fills up to max_out launch-site records in first-
 * launch order and sets *count to the number of sites. */
int fipi_profile_steps(fipi_ctx *ctx, int32_t nsteps, fipi_kernel_stat *out, int32_t max_out, int32_t *count);

/* ---- unit entry points (device pointers, context stream, synchronous) -- */
/* 3-D real-to-complex forward transform, unnormalised (FFTW_FORWARD sign),
 * of one real field into the half spectrum [nz][ny][nx/2+1] complex
 * (interleaved re,im).  Replaces field::fft (field.cpp:438-443) for a real
 * field.  Element type: double for FIPI_FP64 contexts, float otherwise. */
int fipi_fft_r2c(fipi_ctx *ctx, const void *d_real, void *d_half);
/* Inverse (complex-to-real) of the half spectrum, normalised by 1/G as
 * field::ifft + normalize (field.cpp:447-455, primitive.cpp:54-62).  The
 * input is not preserved. */
int fipi_fft_c2r(fipi_ctx *ctx, void *d_half, void *d_real);

/* Cell list of n particles (d_pos: double[n][3]) on the contact grid,
 * replaces cluster::sorting (cluster.cpp:116-169): d_cell_start[ncell+1],
 * d_ids[n] with ids ascending in each cell.  *ncell_out gets ncell. */
int fipi_cells_build(fipi_ctx *ctx, const double *d_pos, int32_t n, int32_t *d_cell_start, int32_t *d_ids,
                     int32_t *ncell_out);
/* The same cell list without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No context: the contact grid comes from params (r_c,
 * margin, box).  Asynchronous on `stream` (a hipStream_t, NULL = the null
 * stream); *d_err (device int) receives the error bits once the work has
 * run: 0, or 1 when a particle lies outside the sorting range (the
 * reference exits there, cluster.cpp:144-163).  *ncell_out (host, may be
 * NULL) gets ncell at once.  The temporaries live in one static device
 * array of at most 16 MiB per GPU, held until the process ends: n must not
 * exceed FIPI_NOALLOC_MAX_PARTICLES and the contact grid
 * FIPI_NOALLOC_MAX_CELLS cells (FIPI_ERR_ARG otherwise).  Every call on a
 * GPU shares that array, so calls that overlap on one GPU race. */
#define FIPI_NOALLOC_MAX_PARTICLES 262144
#define FIPI_NOALLOC_MAX_CELLS 262144
int fipi_cells_build_noalloc(const fipi_params *params, const double *d_pos, int32_t n, int32_t *d_cell_start,
                             int32_t *d_ids, int32_t *d_err, int32_t *ncell_out, void *stream);
/* Contact pairs (i<j, |d| < r_eq) of the 27-cell search of
 * cluster::compute_f_pp (cluster.cpp:193-262), sorted by (i, j).  Writes up
 * to max_pairs pairs to d_pairs[2*max_pairs]; *count gets the total. */
int fipi_contact_pairs(fipi_ctx *ctx, const double *d_pos, int32_t n, int32_t *d_pairs, int64_t max_pairs,
                       int64_t *count);
/* Contact forces f_pp (double[n][3]) at ramp time t, cluster.cpp:174-275. */
int fipi_contact_forces(fipi_ctx *ctx, const double *d_pos, int32_t n, int32_t t, double *d_f_pp);
/* Trilinear interpolation of a real grid field (element type as
 * fipi_fft_r2c) to n particles, field::interpolate (field.cpp:223-248). */
int fipi_interp(fipi_ctx *ctx, const void *d_grid, const double *d_pos, int32_t n, double *d_out);
/* Spread n particle forces (double[n][3]) divided by dx*dy*dz onto a
 * 3-component real grid (element type as fipi_fft_r2c), overwriting it,
 * cluster::compute_f_r (cluster.cpp:344-410).  Deterministic: per grid node
 * the contributions are summed in ascending particle order. */
int fipi_spread(fipi_ctx *ctx, const double *d_pos, const double *d_force, int32_t n, void *d_grid3);
/* field::interpolate2 (field.cpp:251-281, kernel size 3) of a real grid
 * (element type as fipi_fft_r2c) to n particles.  As the reference writes
 * it, the nodes of each axis are floor(x/dx)+1 and floor(x/dx)+2 with
 * weights 1-m', m', m' = (m+1)/3 for the fractional position m. */
int fipi_interp2(fipi_ctx *ctx, const void *d_grid, const double *d_pos, int32_t n, double *d_out);
/* element::get_f_rand (element.cpp:109-144) of n particles at step t
 * (0-based) and Heun stage 1 or 2 with the context's kt and noise_seed:
 * d_out[3n] = sqrt(12 pi r_p mu kt/dt) g, g = sqrt(-2 log r1) cos(2 pi r2)
 * redrawn while |g| > 5, r1, r2 in {0, .., 999}/1000.  rand() % 1000 is
 * replaced by a counter-based draw: attempt k (< 64; g = 0 after that) of
 * component d of particle a takes
 *   key = h(h(h(seed) ^ t) ^ ((3 (stage-1) + d) << 32 | a)),
 *   u = h(key + (k+1) * 0xD1B54A32D192ED03),
 *   r1 = (u >> 32) % 1000, r2 = (u mod 2^32) % 1000,
 * h the splitmix64 finaliser.  The step draws the same values. */
int fipi_brownian_forces(fipi_ctx *ctx, int32_t t, int32_t stage, int32_t n, double *d_out);
/* The particle quantities of the physics above after the last Heun stage
 * (host arrays [3n], any may be NULL): uf, f_rand, and f_drag =
 * -6 pi r mu (u - uf) of element::get_f_drag (element.cpp:147-151). */
int fipi_get_particle_fluid(fipi_ctx *ctx, double *uf, double *f_rand, double *f_drag);

/* ---- device fields of the class facade --------------------------------- */
/* The reference's `field` holds fftw_malloc'ed complex arrays and an
 * in_fspace flag (src/field.hpp:35-93, field.cpp:41-46); the facade
 * (fipi_amd/facade/fipi.hpp) keeps its fields in device buffers of the
 * context and runs the class's operators here.  A field is `dim`
 * components (1, 3 or 9) of G = nx*ny*nz complex values (interleaved re,
 * im), component-major.  Buffers are zero-filled; they are the caller's to
 * destroy (fftw_free, field.cpp:458-462).  These entries need a single-box
 * FP64 context and are synchronous. */
int fipi_buffer_create(fipi_ctx *ctx, int64_t bytes, void **d_ptr);
int fipi_buffer_destroy(fipi_ctx *ctx, void *d_ptr);
int fipi_buffer_upload(fipi_ctx *ctx, void *d_dst, const void *h_src, int64_t bytes);
int fipi_buffer_download(fipi_ctx *ctx, void *h_dst, const void *d_src, int64_t bytes);
/* field::fft / field::ifft (field.cpp:438-455): the complex 3-D DFT of
 * each component in place, sign -1 forward unnormalised, +1 backward
 * normalised by 1/G (primitive.cpp:54-62), as two real transforms of the
 * real and imaginary parts. */
int fipi_field_fft(fipi_ctx *ctx, double *d_v, int32_t dim, int32_t sign);
/* The field operators of src/field.cpp with their space rules (complex
 * products when an operand is in K space, imaginary parts zeroed by
 * real-space operators, division by zero giving zero).  In-place
 * operators act on d_out, whose space is a_kspace:
 *   ADD, SUB       out +=/-= b           field.cpp:525-640 (same dim, space)
 *   ADD_S, SUB_S   out +=/-= s           :574-662 (real space only)
 *   MUL_S, DIV_S   out *=,/= s           :711-830
 *   MULEQ, DIVEQ   out *=,/= scalar b    :666-786 (b real space for /=)
 * Others write d_out from the operands:
 *   MUL, DIV       out = a * b, a / b    :946-1046 (component (d1+1)(d2+1)-1)
 *   DOT            out = a . b           :364-405
 *   POWER          out = sum_d a_d^s     :408-431 (real space)
 *   GRAD           out = i k a, zero on each axis' Nyquist plane; b = the
 *                  wave-vector field     :288-328
 *   COPY           out = a               :836-866
 *   RESET          out = 0               :179-193
 *   REAL           real parts of a into d_out as doubles (dim * G)
 *   FROM_REAL      d_a as doubles (out_dim * G) into a real-space field */
#define FIPI_OP_ADD 0
#define FIPI_OP_SUB 1
#define FIPI_OP_ADD_S 2
#define FIPI_OP_SUB_S 3
#define FIPI_OP_MUL_S 4
#define FIPI_OP_DIV_S 5
#define FIPI_OP_MULEQ 6
#define FIPI_OP_DIVEQ 7
#define FIPI_OP_MUL 8
#define FIPI_OP_DIV 9
#define FIPI_OP_DOT 10
#define FIPI_OP_POWER 11
#define FIPI_OP_GRAD 12
#define FIPI_OP_COPY 13
#define FIPI_OP_RESET 14
#define FIPI_OP_REAL 15
#define FIPI_OP_FROM_REAL 16
int fipi_field_op(fipi_ctx *ctx, int32_t op, double *d_out, int32_t out_dim, const double *d_a, int32_t a_dim,
                  int32_t a_kspace, const double *d_b, int32_t b_dim, int32_t b_kspace, double s);
/* element::get_f_pi (element.cpp:75-97) for n particles: local phase field
 * c_loc[n], gradient dcdx[3n], radius[n] and cos(contact angle)[n] (either
 * may be NULL: r_p, neutral wetting) -> f_pi[3n]. */
int fipi_capillary_forces(fipi_ctx *ctx, const double *d_c_loc, const double *d_dcdx, const double *d_radius,
                          const double *d_cos_theta, int32_t n, double *d_f_pi);
/* element::advect (element.cpp:156-173; uf = 0, f_ext = 0) for stage 1 or
 * 2, element::enforce_bc (element.cpp:177-186) for stage 0 (only d_pos is
 * used); vectors [3n]. */
int fipi_advect(fipi_ctx *ctx, int32_t stage, int32_t n, double *d_pos, double *d_pos_old, double *d_u, double *d_u_old,
                const double *d_f_pi, const double *d_f_pp, const double *d_radius);

/* ---- snapshot / restart ------------------------------------------------ */
/* The whole state of a context (fields, the SBDF2 history and ring index,
 * the FP64 mass mode, particles, step counters) as one host buffer: a
 * header (magic "FIPISNAP", version, grid, particle count, precision, t,
 * the y-pass FFT plan of fipi_set_tuning) followed by the raw device
 * arrays.  fipi_snapshot_load restores it into a context created with the
 * same parameters and takes over the plan (tile widths are kept; they do
 * not change results); stepping then continues bit for bit.  The reference writes ASCII VTK output without the history
 * (field.cpp:469-518, cluster.cpp:415-456), which cannot restart a run. */
int64_t fipi_snapshot_size(fipi_ctx *ctx);
int fipi_snapshot_save(fipi_ctx *ctx, void *buf, int64_t size);
int fipi_snapshot_load(fipi_ctx *ctx, const void *buf, int64_t size);

/* ---- tuning: tile widths and FFT plans --------------------------------- */
/* No counterpart in the reference (FFTW plans, field.cpp:24-139, are made
 * once with FFTW_ESTIMATE).  The launch choices of a context: tile widths
 * (64 or 128 bytes) of the y passes, the CH hub and the Stokes hub, and the
 * FFT plan of the y passes as values per thread (first/last radix: 256
 * points 8.4.8 or 4.4.4.4 in FP64, 16.16 or 8.4.8 in FP32, 128 points 8.2.8
 * or 4.8.4, 64 points 8.8 or 4.4.4;
 * other sizes have one plan; 0 in fipi_set_tuning = the default).  Tile
 * widths do not change results; the plan changes rounding only.
 * fipi_tune times `steps` steps with every combination, keeps the fastest
 * and restores the state from before the trials (kept in device memory
 * meanwhile; about the size of a snapshot). */
typedef struct fipi_tuning_t {
    int32_t y_tile_bytes, hub_tile_bytes, stokes_tile_bytes;
    int32_t y_values_per_thread;
} fipi_tuning_t;
int fipi_tune(fipi_ctx *ctx, int32_t steps);
int fipi_get_tuning(const fipi_ctx *ctx, fipi_tuning_t *out);
int fipi_set_tuning(fipi_ctx *ctx, const fipi_tuning_t *in);

/* ---- batched ensemble of independent boxes ---------------------------- */
/* No counterpart in the reference, which runs one box per process
 * (src/main.cpp:19-252).  Many independent periodic cubic boxes of mixed
 * edge lengths advance together: boxes of equal edge share one slab, every
 * pass is one launch over all boxes of a slab, and a whole ensemble step is
 * one replay of a captured HIP graph (after the first, first-order step).
 * Physics parameters are common; each box has its own state. */
typedef struct fipi_ensemble fipi_ensemble;
typedef struct fipi_box_desc {
    int32_t n;            /* cells per edge (16, 32, 48, 64, 80, 96, 128, 256) */
    int32_t n_particles;  /* spheres in this box */
} fipi_box_desc;
/* common: physics, precision and switches; its grid spacing lx/nx is the
 * spacing of every box (box edge = n * lx/nx); nx..nz and n_particles are
 * ignored.  Boxes are numbered 0..n_boxes-1 in the order given. */
int fipi_ensemble_create(const fipi_params *common, int32_t n_boxes, const fipi_box_desc *boxes,
                         fipi_ensemble **out);
int fipi_ensemble_destroy(fipi_ensemble *ens);
const char *fipi_ensemble_last_error(const fipi_ensemble *ens);
/* Initial state of one box (host arrays: phi[n^3], pos[3*n_particles],
 * radius and cos_theta may be NULL).  Every box must be set before the
 * first step; setting a box again restarts the whole ensemble at t = 0. */
int fipi_ensemble_set_box(fipi_ensemble *ens, int32_t box, const double *phi, const double *pos,
                          const double *radius, const double *cos_theta);
/* Pick the tile widths and the y-pass FFT plan (fipi_tuning_t) of each edge
 * length by timing `steps` steps per candidate, then restart from the
 * initial state.  Tile widths do not change results; the plan changes
 * rounding only. */
int fipi_ensemble_tune(fipi_ensemble *ens, int32_t steps);
/* Advance every box nsteps steps (one graph replay per step); _async
 * leaves the error check to fipi_ensemble_sync. */
int fipi_ensemble_step(fipi_ensemble *ens, int32_t nsteps);
int fipi_ensemble_step_async(fipi_ensemble *ens, int32_t nsteps);
int fipi_ensemble_sync(fipi_ensemble *ens);
void *fipi_ensemble_stream(fipi_ensemble *ens);
int fipi_ensemble_get_phi(fipi_ensemble *ens, int32_t box, double *phi);
int fipi_ensemble_get_particles(fipi_ensemble *ens, int32_t box, double *pos);
int fipi_ensemble_diagnostics(fipi_ensemble *ens, int32_t box, fipi_diag *out);
/* Tile widths chosen for edge n (bytes; 0 when no box has that edge). */
/* the choices of the slab of edge n (all zero when there is none); setting
 * them drops the captured step graphs (recaptured at the next step) */
int fipi_ensemble_get_tuning(const fipi_ensemble *ens, int32_t n, fipi_tuning_t *out);
int fipi_ensemble_set_tuning(fipi_ensemble *ens, int32_t n, const fipi_tuning_t *in);

/* ---- adjoint of the step for calibration ------------------------------- */
/* No counterpart in the reference.  Fits mobility, lambda and per-particle
 * contact angles to a target field: the particles are held at their
 * positions and wet the fluid through a source in the chemical potential,
 *   f = (lambda/eps^2)(c^3 - c) + W,  W = -a_w sum_p cos(theta_p) S_p,
 * S_p the trilinear spreading kernel of particle p over dx*dy*dz (W = 0 at
 * theta = 90 deg or a_w = 0, where the step is the reference step).  The
 * loss is sum_x (phi_T - phi_target)^2 after `steps` steps from phi0.
 * Gradients come from a hand-written adjoint of the step with the forward
 * state checkpointed every checkpoint_every steps and each segment
 * recomputed.  Requires rho1 == rho2. */
typedef struct fipi_adjoint fipi_adjoint;
int fipi_adjoint_create(const fipi_params *params, const double *phi0, const double *pos, double wetting,
                        int32_t steps, int32_t checkpoint_every, fipi_adjoint **out);
int fipi_adjoint_destroy(fipi_adjoint *adj);
const char *fipi_adjoint_last_error(const fipi_adjoint *adj);
int fipi_adjoint_set_target(fipi_adjoint *adj, const double *phi_target);
/* Forward run at (mobility, lambda, theta[n] in radians): the loss (if a
 * target is set and loss != NULL) and the final field (if phi_final). */
int fipi_adjoint_forward(fipi_adjoint *adj, double mobility, double lambda, const double *theta, double *loss,
                         double *phi_final);
/* Loss and dloss/dmobility, dloss/dlambda, dloss/dtheta[n]. */
int fipi_adjoint_gradient(fipi_adjoint *adj, double mobility, double lambda, const double *theta, double *loss,
                          double *g_mobility, double *g_lambda, double *g_theta);

/* ---- seeded synthetic inputs (host only, no GPU needed) ---------------- */
/* phi[i] = mean + U(-amp, amp) from a counter-based generator. */
int fipi_init_phi(int64_t g, double mean, double amp, uint64_t seed, double *phi);
/* Random sequential addition of up to n spheres of radius r with minimum
 * surface gap `gap` in the periodic box (lx, ly, lz); positions written to
 * pos[3n].  Returns the number placed (< n when the box jams within
 * max_attempts tries per sphere), or a negative status. */
int64_t fipi_init_rsa(int64_t n, double r, double gap, double lx, double ly, double lz, uint64_t seed,
                      int64_t max_attempts, double *pos);
/* out[i] = U(lo, hi) from a counter-based generator (e.g. contact angles). */
int fipi_init_uniform(int64_t n, double lo, double hi, uint64_t seed, double *out);

/* ---- the same seeded inputs generated on the GPU ------------------------ */
/* Device versions of the three generators above: the same values, written
 * to device memory on `stream` (a hipStream_t, NULL = the null stream);
 * they return when the values are written.  They replace the reference's
 * unseeded initialisation (main.cpp:24, field.cpp:85-99, cluster.cpp:53-113)
 * for states too large to build on the host.  fipi_init_rsa_device places
 * the spheres by chunks of candidates checked in parallel and returns the
 * same positions as fipi_init_rsa (or a negative status; the message is
 * fipi_init_last_error()). */
int fipi_init_phi_device(int64_t g, double mean, double amp, uint64_t seed, double *d_phi, void *stream);
int fipi_init_uniform_device(int64_t n, double lo, double hi, uint64_t seed, double *d_out, void *stream);
int64_t fipi_init_rsa_device(int64_t n, double r, double gap, double lx, double ly, double lz, uint64_t seed,
                             int64_t max_attempts, double *d_pos, void *stream);
const char *fipi_init_last_error(void);
/* fipi_set_state from device arrays (phi: G doubles, pos: 3n, radius and
 * cos_theta: n doubles or NULL); the state equals what fipi_set_state makes
 * from the same values.  Single-box contexts. */
int fipi_set_state_device(fipi_ctx *ctx, const double *d_phi, const double *d_pos, const double *d_radius,
                          const double *d_cos_theta);

#ifdef __cplusplus
}
#endif
#endif
