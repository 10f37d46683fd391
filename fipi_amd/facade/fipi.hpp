// C++ facade of the MI355X FIPI step, keeping the class interfaces the
// reference driver uses (src/field.hpp:21-93, src/cluster.hpp:25-92,
// src/element.hpp:21-93, src/primitive.hpp:37-81) so that a main() written
// like src/main.cpp drops the GPU step in for its time-loop body
// (main.cpp:75-222).
//
// Header only; links against libfipi.so through the C ABI of
// include/fipi.h (no HIP headers needed).  Differences from the reference:
//   * parameters are runtime values (fipi::config), not head.hpp constants;
//   * errors throw fipi::error carrying the ABI status, instead of exit(0);
//   * fields live on the host only for setup and output; the step itself
//     runs on the device through fipi::solver;
//   * no FFTW plan arguments: fft()/ifft() are not part of the facade, the
//     solver owns its transforms.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fipi.h"

namespace fipi {

// ---------------------------------------------------------------------------
// configuration (the role of src/head.hpp / src/parameter.hpp)
// ---------------------------------------------------------------------------
// The reference's constants (head.hpp:27-74, parameter.hpp:18-20) for
// n_particles spheres and a run of t_end steps.
inline fipi_params reference_params(int n_particles = 20000, int t_end = 20000)
{
    fipi_params p{};
    const double pi = 3.1415926;
    p.nx = p.ny = p.nz = 64;
    p.lx = p.ly = p.lz = 2 * pi;
    p.pi = pi;
    const double dx = p.lx / p.nx;
    p.dt = 1e-3;
    p.ramp_t = 0.2 * t_end;
    p.sigma = 0.07;
    p.epsilon = 3.0 * dx / std::pow(2, 0.5) / 2.0;
    p.lambda = p.sigma * p.epsilon * 3.0 / 2.0 / std::pow(2, 0.5);
    const double r_drop = 0.4 * p.lx;
    p.gravity = -1e1;
    p.rho1 = 1e-3;
    const double Bo = 0.0;
    p.rho2 = p.rho1 - Bo * p.sigma / p.gravity / r_drop / r_drop;
    p.mu = 1e-3;
    p.a_pi = 2;
    p.mobility = std::pow(0.2 * p.epsilon, 2) / p.mu;
    p.r_p = 0.03 * r_drop;
    p.n_particles = n_particles;
    p.r_eq = 5 * p.r_p;
    p.r_c = 5 * p.r_p;
    p.k_n = 0.5 * p.sigma;
    p.margin = 0.0;
    p.tanh2 = std::tanh(2.0);
    p.switch_particle = n_particles > 0;
    p.switch_ch = 1;
    p.switch_ns = 1;
    p.precision = FIPI_FP64;
    return p;
}

// Benchmark units: dx = 1, eps = 1, M = 1, kappa = 1, dt = 0.01, eta = 1,
// sigma = 2 sqrt(2) kappa / (3 eps), r_eq = 2 r_p, contact cells of `cell`.
inline fipi_params benchmark_params(int n, int n_particles, double r_p, int t_end, int precision = FIPI_FP64,
                                    double cell = 6.0)
{
    fipi_params p{};
    p.nx = p.ny = p.nz = n;
    p.lx = p.ly = p.lz = (double)n;
    p.pi = 3.1415926;
    p.dt = 0.01;
    p.ramp_t = 0.2 * t_end;
    p.epsilon = 1.0;
    p.lambda = 1.0;
    p.mobility = 1.0;
    p.mu = 1.0;
    p.sigma = 2.0 * std::sqrt(2.0) * p.lambda / (3.0 * p.epsilon);
    p.a_pi = 2.0;
    p.k_n = 0.5 * p.sigma;
    p.r_p = r_p;
    p.r_eq = 2.0 * r_p;
    p.r_c = cell;
    p.rho1 = p.rho2 = 1e-3;
    p.gravity = -10.0;
    p.tanh2 = std::tanh(2.0);
    p.n_particles = n_particles;
    p.switch_particle = n_particles > 0;
    p.switch_ch = 1;
    p.switch_ns = 1;
    p.precision = precision;
    return p;
}

struct config {
    static fipi_params &get()
    {
        static fipi_params p = reference_params();
        return p;
    }
    static void set(const fipi_params &p) { get() = p; }
};

class error : public std::runtime_error {
  public:
    int code;
    error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const fipi_ctx *ctx = nullptr)
{
    if (rc != FIPI_OK) throw error(rc, std::string("fipi: ") + fipi_last_error(ctx));
}

// ---------------------------------------------------------------------------
// primitives (src/primitive.hpp:18-81)
// ---------------------------------------------------------------------------
class vector {
    double x = 0, y = 0, z = 0;

  public:
    vector() = default;
    vector(double a, double b, double c) : x(a), y(b), z(c) {}
    void set_value(double a, double b, double c) { x = a; y = b; z = c; }
    double get_x() const { return x; }
    double get_y() const { return y; }
    double get_z() const { return z; }
    double norm() const { return std::sqrt((x * x + y * y) + z * z); }
};

// output_time, primitive.cpp:21-25: every t_end/100 steps
inline bool output_time(int t, int t_end) { return std::remainder((t + 1) * 100.0, (double)t_end) == 0; }

// convert_to_vtk, primitive.cpp:28-36
inline std::string convert_to_vtk(int n, const std::string &name)
{
    std::stringstream s;
    s << "data/" << name << n << ".vtk";
    return s.str();
}

// ---------------------------------------------------------------------------
// field (src/field.hpp:21-93): host mirror of a grid field
// ---------------------------------------------------------------------------
class field {
    std::string type;
    int dim = 1;
    int nx, ny, nz;
    std::vector<double> v;  // dim * G reals, component-major

  public:
    explicit field(const std::string &type_id) : type(type_id)
    {
        const fipi_params &p = config::get();
        nx = p.nx;
        ny = p.ny;
        nz = p.nz;
        if (type == "SCALARS") dim = 1;
        else if (type == "VECTORS") dim = 3;
        else if (type == "TENSORS") dim = 9;
        else throw error(FIPI_ERR_ARG, "UNDEFINED FIELD TYPE");  // field.cpp:35-38
        v.assign((size_t)dim * size(), 0.0);
    }

    // named constructors, field.cpp:62-139: "phase_field" gets
    // mean + U(-amp, amp) from a seed, "wave_vector" the reference wave
    // numbers, "gravity" (0, 0, gravity)
    field(const std::string &type_id, const std::string &name_id, uint64_t seed = 1, double mean = 0.0,
          double amp = 0.1)
        : field(type_id)
    {
        const fipi_params &p = config::get();
        if (name_id == "phase_field") {
            check(fipi_init_phi((int64_t)size(), mean, amp, seed, v.data()));
        } else if (name_id == "wave_vector") {
            for (size_t it = 0; it < size(); it++) {
                const int k = (int)(it / ((size_t)nx * ny)), j = (int)(it % ((size_t)nx * ny) / nx),
                          i = (int)(it % ((size_t)nx * ny) % nx);
                v[it] = (i <= nx / 2 ? i : i - nx) * 2.0 * p.pi / p.lx;
                v[size() + it] = (j <= ny / 2 ? j : j - ny) * 2.0 * p.pi / p.ly;
                v[2 * size() + it] = (k <= nz / 2 ? k : k - nz) * 2.0 * p.pi / p.lz;
            }
        } else if (name_id == "gravity") {
            for (size_t it = 0; it < size(); it++) v[2 * size() + it] = p.gravity;
        } else {
            throw error(FIPI_ERR_ARG, "UNDEFINED FIELD NAME");  // field.cpp:134-138
        }
    }

    size_t size() const { return (size_t)nx * ny * nz; }
    int dimension() const { return dim; }
    double *data(int d = 0) { return v.data() + (size_t)d * size(); }
    const double *data(int d = 0) const { return v.data() + (size_t)d * size(); }

    void set_value(int d, int it, double value) { v[(size_t)d * size() + it] = value; }   // field.cpp:142-147
    void add_value(int d, int it, double value) { v[(size_t)d * size() + it] += value; }  // field.cpp:150-155
    void reset(const std::string &) { std::fill(v.begin(), v.end(), 0.0); }                  // field.cpp:179-193

    // field::interpolate, field.cpp:223-248 (host; the solver interpolates on the device)
    double interpolate(const vector &pos, int d) const
    {
        const fipi_params &p = config::get();
        const double dx = p.lx / nx, dy = p.ly / ny, dz = p.lz / nz;
        const int i = (int)std::floor(pos.get_x() / dx), j = (int)std::floor(pos.get_y() / dy),
                  k = (int)std::floor(pos.get_z() / dz);
        const double m = pos.get_x() / dx - std::floor(pos.get_x() / dx);
        const double n = pos.get_y() / dy - std::floor(pos.get_y() / dy);
        const double o = pos.get_z() / dz - std::floor(pos.get_z() / dz);
        const double *a = data(d);
        auto at = [&](int ii, int jj, int kk) { return a[ii + (size_t)jj * nx + (size_t)kk * nx * ny]; };
        const int i1 = (i + 1) % nx, j1 = (j + 1) % ny, k1 = (k + 1) % nz;
        double r = (1.0 - m) * (1.0 - n) * (1.0 - o) * at(i, j, k);
        r += (1.0 - m) * n * (1.0 - o) * at(i, j1, k);
        r += m * (1.0 - n) * (1.0 - o) * at(i1, j, k);
        r += m * n * (1.0 - o) * at(i1, j1, k);
        r += (1.0 - m) * (1.0 - n) * o * at(i, j, k1);
        r += (1.0 - m) * n * o * at(i, j1, k1);
        r += m * (1.0 - n) * o * at(i1, j, k1);
        r += m * n * o * at(i1, j1, k1);
        return r;
    }

    // field::output_to_file, field.cpp:469-518: legacy ASCII VTK
    void output_to_file(const std::string &file_name, const std::string &variable_name) const
    {
        std::ofstream out(file_name.c_str());
        out << "# vtk DataFile Version 2.0\n" << variable_name << " PROFILE\nASCII\n\n";
        out << "DATASET STRUCTURED_POINTS\nDIMENSIONS " << nx << " " << ny << " " << nz << "\n";
        out << "ASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA " << size() << "\n";
        out << type << " " << variable_name << " double\n";
        if (dim == 1) out << "LOOKUP_TABLE default\n";
        for (size_t it = 0; it < size(); it++)
            for (int d = 0; d < dim; d++) out << data(d)[it] << (d == dim - 1 ? "\n" : " ");
    }
};

// ---------------------------------------------------------------------------
// element / cluster (src/element.hpp:21-93, src/cluster.hpp:25-92)
// ---------------------------------------------------------------------------
// Per-particle state in structure-of-arrays form; element(i) gives the view
// the reference's element class holds (element.hpp:28-58).
class cluster {
  public:
    int N = 0;
    std::vector<double> p, p_old, u, u_old, f_pp, f_pi;  // 3N each
    std::vector<double> radius, cos_theta;               // N each
    int nx_cell = 0, ny_cell = 0, nz_cell = 0;
    double dx_cell = 0, dy_cell = 0, dz_cell = 0;
    double tau = 0.0;
    int it_mark = -1;

    // cluster.cpp:21-40
    cluster()
    {
        const fipi_params &q = config::get();
        N = q.switch_particle ? q.n_particles : 0;
        nx_cell = (int)std::floor(q.lx / q.r_c + 2 * q.margin / q.r_c);
        ny_cell = (int)std::floor(q.ly / q.r_c + 2 * q.margin / q.r_c);
        nz_cell = (int)std::floor(q.lz / q.r_c + 2 * q.margin / q.r_c);
        dx_cell = std::fmod(q.lx + 2 * q.margin, q.r_c) / nx_cell + q.r_c;
        dy_cell = std::fmod(q.ly + 2 * q.margin, q.r_c) / ny_cell + q.r_c;
        dz_cell = std::fmod(q.lz + 2 * q.margin, q.r_c) / nz_cell + q.r_c;
        for (auto *a : {&p, &p_old, &u, &u_old, &f_pp, &f_pi}) a->assign(3 * (size_t)N, 0.0);
        radius.assign(N, q.r_p);
        cos_theta.assign(N, 0.0);
    }

    vector position(int i) const { return vector(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

    // cluster::distribute, cluster.cpp:53-92, as seeded random sequential
    // addition with a minimum surface gap.  Throws if the box jams.
    void distribute(uint64_t seed = 2, double gap = 0.0)
    {
        const fipi_params &q = config::get();
        const int64_t got = fipi_init_rsa(N, q.r_p, gap, q.lx, q.ly, q.lz, seed, 100000, p.data());
        if (got < 0) check((int)got);
        if (got != N) throw error(FIPI_ERR_ARG, "random sequential addition jammed before placing all spheres");
        p_old = p;  // cluster.cpp:88
    }

    // per-particle contact angles (cos theta = 0 is the reference's neutral wetting)
    void set_contact_angles_deg(double lo, double hi, uint64_t seed = 3)
    {
        std::vector<double> th(N);
        check(fipi_init_uniform(N, lo, hi, seed, th.data()));
        const double pi = 3.14159265358979323846;
        for (int i = 0; i < N; i++) cos_theta[i] = std::cos(th[i] * pi / 180.0);
    }

    // cluster::summary, cluster.cpp:460-482
    void summary() const
    {
        const fipi_params &q = config::get();
        const double t_ch = std::pow(q.epsilon, 3) / q.mobility / q.sigma;
        std::cout << "----------------SUMMARY------------------\n";
        std::cout << "Delta t = " << q.dt << "\n\n";
        std::cout << "CH time scale = " << t_ch << "\n";
        std::cout << "S = " << std::sqrt(q.mu * q.mobility) / q.epsilon << "\n";
        std::cout << "l_map/r_p = " << 2 * std::sqrt(2.0) * q.epsilon / q.r_p << "\n\n";
        std::cout << "------------------------------------------\n";
        std::cout << "Fluid viscosity mu = " << q.mu << "\n";
        std::cout << "number of particles = " << N << "\n";
        std::cout << "k_n_initial = " << q.k_n << "\n\n";
    }

    // cluster::output, cluster.cpp:415-456: legacy ASCII VTK of the particles
    void output(const std::string &name) const
    {
        const fipi_params &q = config::get();
        const double dx = q.lx / q.nx;
        std::ofstream tr(name.c_str());
        tr << "# vtk DataFile Version 2.0\nparticles\nASCII\n\nDATASET UNSTRUCTURED_GRID\n";
        tr << "POINTS " << N << " double\n";
        for (int i = 0; i < N; i++) tr << p[3 * i] / dx << " " << p[3 * i + 1] / dx << " " << p[3 * i + 2] / dx << "\n";
        tr << "\nPOINT_DATA " << N << "\nSCALARS radius double\nLOOKUP_TABLE default\n";
        for (int i = 0; i < N; i++) tr << radius[i] / dx << "\n";
        tr << "\nSCALARS f_pp double\nLOOKUP_TABLE default\n";
        for (int i = 0; i < N; i++) {
            const vector f(f_pp[3 * i], f_pp[3 * i + 1], f_pp[3 * i + 2]);
            tr << f.norm() / q.a_pi / q.pi / q.sigma / q.r_p << "\n";
        }
        std::cout << "packing energy= " << tau << " ";
    }
};

// ---------------------------------------------------------------------------
// solver: the reference time-loop body (main.cpp:75-222) on the GPU
// ---------------------------------------------------------------------------
class solver {
    fipi_ctx *ctx = nullptr;

  public:
    solver() { check(fipi_create(&config::get(), &ctx)); }
    ~solver() { fipi_destroy(ctx); }
    solver(const solver &) = delete;
    solver &operator=(const solver &) = delete;

    // initial state: phi from `c`, particles from `s`; velocity zero
    void load(const field &c, const cluster &s)
    {
        check(fipi_set_state(ctx, c.data(), s.N ? s.p.data() : nullptr, s.N ? s.radius.data() : nullptr,
                             s.N ? s.cos_theta.data() : nullptr),
              ctx);
    }

    void step(int nsteps = 1) { check(fipi_step(ctx, nsteps), ctx); }

    int time()
    {
        int32_t t = 0;
        check(fipi_get_time(ctx, &t), ctx);
        return t;
    }

    // copy the device state back into the host mirrors (any may be null)
    void fetch(field *c, field *u = nullptr, field *grad_c = nullptr, cluster *s = nullptr)
    {
        if (c) check(fipi_get_phi(ctx, c->data()), ctx);
        if (u) check(fipi_get_velocity(ctx, u->data()), ctx);
        if (grad_c) check(fipi_get_grad_phi(ctx, grad_c->data()), ctx);
        if (s && s->N) {
            check(fipi_get_particles(ctx, s->p.data(), s->p_old.data(), s->u.data(), s->u_old.data(), s->f_pp.data(),
                                     s->f_pi.data()),
                  ctx);
            fipi_diag d{};
            check(fipi_diagnostics(ctx, &d), ctx);
            s->tau = d.tau;
            s->it_mark = d.it_mark;
        }
    }

    fipi_diag diagnostics()
    {
        fipi_diag d{};
        check(fipi_diagnostics(ctx, &d), ctx);
        return d;
    }

    fipi_ctx *handle() { return ctx; }
};

}  // namespace fipi
