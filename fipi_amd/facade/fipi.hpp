// C++ facade of the MI355X FIPI step, keeping the class interfaces the
// reference driver uses (src/field.hpp:21-93, src/cluster.hpp:25-92,
// src/element.hpp:21-93, src/primitive.hpp:37-81) so that a main() written
// like src/main.cpp drops the GPU step in for its time-loop body
// (main.cpp:75-222).
//
// Header only; links against libfipi.so through the C ABI of
// include/fipi.h (no HIP headers needed).  Two ways to drive a run:
//   * fipi::solver::step() replaces the whole loop body of main.cpp
//     (main.cpp:75-222) with the fused GPU step;
//   * the classes below keep the reference's methods, so the loop body can
//     also be written statement by statement as main.cpp has it: `field`
//     values live in device buffers and their operators, transforms,
//     gradients and products run in libfipi (fipi_field_*); `cluster`
//     calls the unit entries (cells, contacts, interpolation, spread,
//     capillary forces, advection).
// Differences from the reference:
//   * parameters are runtime values (fipi::config), not head.hpp constants;
//   * errors throw fipi::error carrying the ABI status, instead of exit(0);
//   * fft()/ifft() take a fipi::plan (the direction) instead of an FFTW plan
//     (field.hpp:55-56); the facade's device context owns the transforms;
//   * fields have proper copy and move semantics (the reference declares no
//     copy constructor and relies on copy elision).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
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
// the facade's device context (fields and unit entries), FP64, made from
// config::get() on first use
// ---------------------------------------------------------------------------
class device {
    static fipi_ctx *&slot()
    {
        static fipi_ctx *ctx = nullptr;
        return ctx;
    }

  public:
    static fipi_ctx *ctx()
    {
        fipi_ctx *&c = slot();
        if (!c) {
            fipi_params p = config::get();
            p.precision = FIPI_FP64;
            check(fipi_create(&p, &c));
        }
        return c;
    }
    // drop the context (after config::set with another grid)
    static void reset()
    {
        fipi_ctx *&c = slot();
        if (c) fipi_destroy(c);
        c = nullptr;
    }
};

// a device buffer of the facade's context
class dbuf {
    void *p = nullptr;
    size_t n = 0;

  public:
    dbuf() = default;
    explicit dbuf(size_t bytes) : n(bytes)
    {
        check(fipi_buffer_create(device::ctx(), (int64_t)bytes, &p), device::ctx());
    }
    ~dbuf()
    {
        if (p) fipi_buffer_destroy(device::ctx(), p);
    }
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    dbuf(dbuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    dbuf &operator=(dbuf &&o) noexcept
    {
        std::swap(p, o.p);
        std::swap(n, o.n);
        return *this;
    }
    double *d() const { return static_cast<double *>(p); }
    size_t bytes() const { return n; }
    void upload(const void *h, size_t bytes) { check(fipi_buffer_upload(device::ctx(), p, h, (int64_t)bytes), device::ctx()); }
    void download(void *h, size_t bytes) const
    {
        check(fipi_buffer_download(device::ctx(), h, p, (int64_t)bytes), device::ctx());
    }
};

// FFTW_FORWARD / FFTW_BACKWARD plans (main.cpp:70-73) become directions
struct plan {
    int sign;
};
inline plan plan_forward() { return plan{-1}; }
inline plan plan_backward() { return plan{+1}; }

// ---------------------------------------------------------------------------
// field (src/field.hpp:21-93): dim components of complex values on the
// device, and the in_fspace flag
// ---------------------------------------------------------------------------
class field {
    std::string type;
    int dim = 1;
    int nx = 0, ny = 0, nz = 0;
    bool in_fspace = false;
    dbuf v;  // dim * G complex (re, im)

    static int dim_of(const std::string &t)
    {
        if (t == "SCALARS") return 1;
        if (t == "VECTORS") return 3;
        if (t == "TENSORS") return 9;
        throw error(FIPI_ERR_ARG, "UNDEFINED FIELD TYPE");  // field.cpp:35-38
    }
    static std::string type_of(int d) { return d == 1 ? "SCALARS" : (d == 3 ? "VECTORS" : "TENSORS"); }
    void op(int code, const field *a, const field *b, double s = 0.0)
    {
        // in-place operators take this field's space as the left operand's
        const bool in_place = code <= FIPI_OP_DIVEQ;
        check(fipi_field_op(device::ctx(), code, v.d(), dim, a ? a->v.d() : nullptr, a ? a->dim : 0,
                            in_place ? (int)in_fspace : (a ? (int)a->in_fspace : 0), b ? b->v.d() : nullptr,
                            b ? b->dim : 0, b ? (int)b->in_fspace : 0, s),
              device::ctx());
    }
    static void fail(const char *msg) { throw error(FIPI_ERR_ARG, msg); }

  public:
    explicit field(const std::string &type_id) : type(type_id), dim(dim_of(type_id))
    {
        const fipi_params &p = config::get();
        nx = p.nx;
        ny = p.ny;
        nz = p.nz;
        v = dbuf(sizeof(double) * 2 * (size_t)dim * size());
    }

    // named fields, field.cpp:62-139: "phase_field" mean + U(-amp, amp)
    // from a seed (the reference uses rand()), "wave_vector" the reference
    // wave numbers (flagged real space, field.cpp:81), "gravity" (0, 0, g)
    field(const std::string &type_id, const std::string &name_id, uint64_t seed = 1, double mean = 0.0,
          double amp = 0.1)
        : field(type_id)
    {
        const fipi_params &p = config::get();
        std::vector<double> h((size_t)dim * size(), 0.0);
        if (name_id == "phase_field") {
            check(fipi_init_phi((int64_t)size(), mean, amp, seed, h.data()));
        } else if (name_id == "wave_vector") {
            if (dim != 3) fail("wave_vector is a VECTORS field");
            for (size_t it = 0; it < size(); it++) {
                const int k = (int)(it / ((size_t)nx * ny)), j = (int)(it % ((size_t)nx * ny) / nx),
                          i = (int)(it % ((size_t)nx * ny) % nx);
                h[it] = (i <= nx / 2 ? i : i - nx) * 2.0 * p.pi / p.lx;
                h[size() + it] = (j <= ny / 2 ? j : j - ny) * 2.0 * p.pi / p.ly;
                h[2 * size() + it] = (k <= nz / 2 ? k : k - nz) * 2.0 * p.pi / p.lz;
            }
        } else if (name_id == "gravity") {
            if (dim != 3) fail("gravity is a VECTORS field");
            for (size_t it = 0; it < size(); it++) h[2 * size() + it] = p.gravity;
        } else {
            throw error(FIPI_ERR_ARG, "UNDEFINED FIELD NAME");  // field.cpp:134-138
        }
        for (int d = 0; d < dim; d++) assign_real(h.data() + (size_t)d * size(), d);
    }

    field(const field &o) : field(o.type)
    {
        in_fspace = o.in_fspace;
        op(FIPI_OP_COPY, &o, nullptr);
    }
    field(field &&o) noexcept = default;
    field &operator=(field &&o) noexcept
    {
        if (o.dim != dim) {
            // as the copy assignment, a field keeps its type
            field &&r = std::move(o);
            return *this = static_cast<const field &>(r);
        }
        type = std::move(o.type);
        in_fspace = o.in_fspace;
        v = std::move(o.v);
        return *this;
    }
    // field.cpp:836-866
    field &operator=(const field &rhs)
    {
        if (this == &rhs) return *this;
        if (dim != rhs.dim) fail("= RHS AND LHS FIELD NOT OF THE SAME TYPE");
        in_fspace = rhs.in_fspace;
        op(FIPI_OP_COPY, &rhs, nullptr);
        return *this;
    }

    size_t size() const { return (size_t)nx * ny * nz; }
    int dimension() const { return dim; }
    bool kspace() const { return in_fspace; }
    double *device_data() const { return v.d(); }

    // real parts of component d on the host, and a real field from the host
    std::vector<double> real(int d = 0) const
    {
        std::vector<double> c(2 * size()), r(size());
        check(fipi_buffer_download(device::ctx(), c.data(), v.d() + 2 * (size_t)d * size(),
                                   (int64_t)(sizeof(double) * 2 * size())),
              device::ctx());
        for (size_t i = 0; i < size(); i++) r[i] = c[2 * i];
        return r;
    }
    void assign_real(const double *h, int d = 0)
    {
        std::vector<double> c(2 * size(), 0.0);
        for (size_t i = 0; i < size(); i++) c[2 * i] = h[i];
        check(fipi_buffer_upload(device::ctx(), v.d() + 2 * (size_t)d * size(), c.data(),
                                 (int64_t)(sizeof(double) * 2 * size())),
              device::ctx());
        in_fspace = false;
    }

    // a real-space field from dim * G doubles in device memory
    void assign_device_real(const double *d_real)
    {
        check(fipi_field_op(device::ctx(), FIPI_OP_FROM_REAL, v.d(), dim, d_real, dim, 0, nullptr, 0, 0, 0.0),
              device::ctx());
        in_fspace = false;
    }

    // field.cpp:142-155: one point, imaginary part zeroed
    void set_value(int d, int it, double value)
    {
        const double c[2] = {value, 0.0};
        check(fipi_buffer_upload(device::ctx(), v.d() + 2 * ((size_t)d * size() + it), c, sizeof(c)), device::ctx());
    }
    void add_value(int d, int it, double value)
    {
        double c[2];
        check(fipi_buffer_download(device::ctx(), c, v.d() + 2 * ((size_t)d * size() + it), sizeof(c)), device::ctx());
        c[0] += value;
        c[1] = 0.0;
        check(fipi_buffer_upload(device::ctx(), v.d() + 2 * ((size_t)d * size() + it), c, sizeof(c)), device::ctx());
    }
    // field.cpp:179-193
    void reset(const std::string &space)
    {
        op(FIPI_OP_RESET, nullptr, nullptr);
        if (space == "f") in_fspace = true;
        else if (space == "p") in_fspace = false;
    }

    // field.cpp:438-455
    void fft(const plan &pl)
    {
        if (pl.sign != -1) fail("fft() takes the forward plan");
        check(fipi_field_fft(device::ctx(), v.d(), dim, -1), device::ctx());
        in_fspace = true;
    }
    void ifft(const plan &pl)
    {
        if (pl.sign != 1) fail("ifft() takes the backward plan");
        check(fipi_field_fft(device::ctx(), v.d(), dim, 1), device::ctx());
        in_fspace = false;
    }

    // vector calculus, field.cpp:288-431
    field grad(const field &kf) const
    {
        if (!in_fspace) fail("YOU CAN ONLY CALCULATE GRADIENT IN FREQUENCY SPACE!");
        if (dim != 1) fail("grad() of a SCALARS field");
        field out("VECTORS");
        out.in_fspace = true;
        out.op(FIPI_OP_GRAD, this, &kf);
        return out;
    }
    field dot(const field &b) const
    {
        if (dim != b.dim) fail("dot() TWO FIELDS WITH DIFFERENT DIMENSIONS");
        field out("SCALARS");
        out.in_fspace = in_fspace || b.in_fspace;
        out.op(FIPI_OP_DOT, this, &b);
        return out;
    }
    field power(double e) const
    {
        if (in_fspace) fail("POWER() FIELD IN FREQUENCY SPACE");
        field out("SCALARS");
        out.op(FIPI_OP_POWER, this, nullptr, e);
        return out;
    }

    // compound assignment, field.cpp:525-830
    field &operator+=(const field &rhs)
    {
        if (dim != rhs.dim) fail("+=FIELD TWO FIELDS WITH DIFFERENT DIMENSIONS");
        if (in_fspace != rhs.in_fspace) fail("+=FIELD TWO FIELDS IN DIFFERENT SPACE");
        op(FIPI_OP_ADD, nullptr, &rhs);
        return *this;
    }
    field &operator-=(const field &rhs)
    {
        if (dim != rhs.dim) fail("-=FIELD TWO FIELDS WITH DIFFERENT DIMENSIONS");
        if (in_fspace != rhs.in_fspace) fail("-=FIELD TWO FIELDS IN DIFFERENT SPACE");
        op(FIPI_OP_SUB, nullptr, &rhs);
        return *this;
    }
    field &operator+=(double rhs)
    {
        if (in_fspace) fail("+=DOUBLE ADDING REAL NUMBER IN FREQUENCY SPACE");
        op(FIPI_OP_ADD_S, nullptr, nullptr, rhs);
        return *this;
    }
    field &operator-=(double rhs)
    {
        if (in_fspace) fail("-=DOUBLE MINUS REAL NUMBER IN FREQUENCY SPACE");
        op(FIPI_OP_SUB_S, nullptr, nullptr, rhs);
        return *this;
    }
    field &operator*=(const field &rhs)
    {
        if (rhs.dim != 1) fail("*= RHS FIELD IS NOT A SCALAR ONE");
        // complex product when either operand is in K space; like the
        // reference, the left operand keeps its space flag
        op(FIPI_OP_MULEQ, nullptr, &rhs);
        return *this;
    }
    field &operator*=(double rhs)
    {
        op(FIPI_OP_MUL_S, nullptr, nullptr, rhs);
        return *this;
    }
    field &operator/=(const field &rhs)
    {
        if (rhs.in_fspace) fail("/= RHS FIELD IN FREQUENCY SPACE");
        if (rhs.dim != 1) fail("/= RHS FIELD IS NOT A SCALAR ONE");
        op(FIPI_OP_DIVEQ, nullptr, &rhs);
        return *this;
    }
    field &operator/=(double rhs)
    {
        op(FIPI_OP_DIV_S, nullptr, nullptr, rhs);
        return *this;
    }

    // binary operators, field.cpp:874-1060
    field operator+(const field &rhs) const { field r(*this); r += rhs; return r; }
    field operator-(const field &rhs) const { field r(*this); r -= rhs; return r; }
    field operator+(double rhs) const { field r(*this); r += rhs; return r; }
    field operator-(double rhs) const { field r(*this); r -= rhs; return r; }
    field operator*(double rhs) const { field r(*this); r *= rhs; return r; }
    field operator/(double rhs) const { field r(*this); r /= rhs; return r; }
    field operator*(const field &rhs) const
    {
        const int od = dim * rhs.dim;
        if (od != 1 && od != 3 && od != 9) fail("* FIELDS OF INCOMPATIBLE TYPES");
        field out(type_of(od));
        out.in_fspace = in_fspace || rhs.in_fspace;
        out.op(FIPI_OP_MUL, this, &rhs);
        return out;
    }
    field operator/(const field &rhs) const
    {
        if (rhs.in_fspace) fail("/FIELD DIVISOR FIELD IN FREQUENCY SPACE");
        const int od = dim * rhs.dim;
        if (od != 1 && od != 3 && od != 9) fail("/ FIELDS OF INCOMPATIBLE TYPES");
        field out(type_of(od));
        out.in_fspace = in_fspace;
        out.op(FIPI_OP_DIV, this, &rhs);
        return out;
    }

    // field::interpolate, field.cpp:223-248, of one point on the host
    double interpolate(const vector &pos, int d) const
    {
        const fipi_params &p = config::get();
        const double dx = p.lx / nx, dy = p.ly / ny, dz = p.lz / nz;
        const int i = (int)std::floor(pos.get_x() / dx), j = (int)std::floor(pos.get_y() / dy),
                  k = (int)std::floor(pos.get_z() / dz);
        const double m = pos.get_x() / dx - std::floor(pos.get_x() / dx);
        const double n = pos.get_y() / dy - std::floor(pos.get_y() / dy);
        const double o = pos.get_z() / dz - std::floor(pos.get_z() / dz);
        const std::vector<double> a = real(d);
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

    // field::output_to_file, field.cpp:469-518: legacy ASCII VTK of the real parts
    void output_to_file(const std::string &file_name, const std::string &variable_name) const
    {
        std::vector<std::vector<double>> c;
        for (int d = 0; d < dim; d++) c.push_back(real(d));
        std::ofstream out(file_name.c_str());
        out << "# vtk DataFile Version 2.0\n" << variable_name << " PROFILE\nASCII\n\n";
        out << "DATASET STRUCTURED_POINTS\nDIMENSIONS " << nx << " " << ny << " " << nz << "\n";
        out << "ASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA " << size() << "\n";
        out << type << " " << variable_name << " double\n";
        if (dim == 1) out << "LOOKUP_TABLE default\n";
        for (size_t it = 0; it < size(); it++)
            for (int d = 0; d < dim; d++) out << c[d][it] << (d == dim - 1 ? "\n" : " ");
    }
};

// non-member operators, field.cpp:1062-1080
inline field operator+(double lhs, const field &rhs) { return rhs + lhs; }
inline field operator*(double lhs, const field &rhs) { return rhs * lhs; }
inline field operator-(double lhs, const field &rhs) { return rhs * (-1.0) + lhs; }

// ---------------------------------------------------------------------------
// node / list (src/linked_list.hpp:20-66): the particle ids of one contact
// cell, appended in ascending id order by cluster::sorting
// ---------------------------------------------------------------------------
class node {
    friend class list;
    int value = 0;
    node *next = nullptr, *last = nullptr;

  public:
    node() = default;
    explicit node(int val) : value(val) {}
    int get_value() const { return value; }
    const node *get_next() const { return next; }
};

class list {
    node *front = nullptr, *end = nullptr;

  public:
    list() = default;
    explicit list(int val) { append(val); }
    ~list() { clear(); }
    list(const list &) = delete;
    list &operator=(const list &) = delete;
    list(list &&o) noexcept : front(o.front), end(o.end) { o.front = o.end = nullptr; }
    void append(int val)
    {
        node *n = new node(val);
        n->last = end;
        if (end) end->next = n;
        else front = n;
        end = n;
    }
    // removes the first node holding val
    void remove(int val)
    {
        for (node *n = front; n; n = n->next)
            if (n->value == val) {
                (n->last ? n->last->next : front) = n->next;
                (n->next ? n->next->last : end) = n->last;
                delete n;
                return;
            }
    }
    void clear()
    {
        for (node *n = front; n;) {
            node *x = n->next;
            delete n;
            n = x;
        }
        front = end = nullptr;
    }
    const node *head() const { return front; }
    std::vector<int> values() const
    {
        std::vector<int> r;
        for (const node *n = front; n; n = n->next) r.push_back(n->value);
        return r;
    }
};

// ---------------------------------------------------------------------------
// element / cluster (src/element.hpp:21-93, src/cluster.hpp:25-92)
// ---------------------------------------------------------------------------
// One particle's state as the reference's element holds it
// (element.hpp:28-58), read from a cluster's host mirror.  f_ext, f_rand,
// f_drag and uf are zero on this path (element.cpp:40-41,100-103; main.cpp
// never calls compute_f_drag / compute_f_rand).
struct element {
    double r = 0.0;
    vector p, p_old, u, u_old;
    vector f_pp, f_pi, f_ext, uf;
    double c = 0.0;
    vector dcdx;
    double volume() const { return 4.0 / 3.0 * config::get().pi * std::pow(r, 3); }
};

// The particles in structure-of-arrays form, kept in device buffers of the
// facade's context (the methods below run there) with a host mirror that
// fetch() / upload() move between the two.
class cluster {
    dbuf dp, dpo, du, duo, dfpp, dfpi, drad, dcth, dc, ddcdx, dstart, dids, grid;
    int ncell = 0;
    bool cells_built = false;

    // real parts of a scalar or vector field in `grid` (3 G doubles)
    const double *grid_of(const field &f)
    {
        if (f.dimension() > 3) throw error(FIPI_ERR_ARG, "interpolation of a TENSORS field");
        check(fipi_field_op(device::ctx(), FIPI_OP_REAL, grid.d(), f.dimension(), f.device_data(), f.dimension(),
                            (int)f.kspace(), nullptr, 0, 0, 0.0),
              device::ctx());
        return grid.d();
    }

  public:
    int N = 0;
    std::vector<double> p, p_old, u, u_old, f_pp, f_pi;  // 3N each (host mirror)
    std::vector<double> radius, cos_theta, c_loc, dcdx;  // N, N, N, 3N
    int nx_cell = 0, ny_cell = 0, nz_cell = 0;
    double dx_cell = 0, dy_cell = 0, dz_cell = 0;
    double tau = 0.0;
    int it_mark = 0;

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
        for (auto *a : {&p, &p_old, &u, &u_old, &f_pp, &f_pi, &dcdx}) a->assign(3 * (size_t)N, 0.0);
        radius.assign(N, q.r_p);
        cos_theta.assign(N, 0.0);
        c_loc.assign(N, 0.0);
        const size_t v = sizeof(double) * 3 * (size_t)(N > 0 ? N : 1), s = sizeof(double) * (size_t)(N > 0 ? N : 1);
        dp = dbuf(v);
        dpo = dbuf(v);
        du = dbuf(v);
        duo = dbuf(v);
        dfpp = dbuf(v);
        dfpi = dbuf(v);
        ddcdx = dbuf(v);
        drad = dbuf(s);
        dcth = dbuf(s);
        dc = dbuf(s);
        dids = dbuf(sizeof(int32_t) * (size_t)(N > 0 ? N : 1));
        dstart = dbuf(sizeof(int32_t) * ((size_t)nx_cell * ny_cell * nz_cell + 1));
        grid = dbuf(sizeof(double) * 3 * (size_t)q.nx * q.ny * q.nz);
        upload();
    }

    // host mirror -> device, device -> host mirror
    void upload()
    {
        if (!N) return;
        const size_t v = sizeof(double) * 3 * (size_t)N, s = sizeof(double) * (size_t)N;
        dp.upload(p.data(), v);
        dpo.upload(p_old.data(), v);
        du.upload(u.data(), v);
        duo.upload(u_old.data(), v);
        dfpp.upload(f_pp.data(), v);
        dfpi.upload(f_pi.data(), v);
        ddcdx.upload(dcdx.data(), v);
        drad.upload(radius.data(), s);
        dcth.upload(cos_theta.data(), s);
        dc.upload(c_loc.data(), s);
    }
    void fetch()
    {
        if (!N) return;
        const size_t v = sizeof(double) * 3 * (size_t)N, s = sizeof(double) * (size_t)N;
        dp.download(p.data(), v);
        dpo.download(p_old.data(), v);
        du.download(u.data(), v);
        duo.download(u_old.data(), v);
        dfpp.download(f_pp.data(), v);
        dfpi.download(f_pi.data(), v);
        ddcdx.download(dcdx.data(), v);
        dc.download(c_loc.data(), s);
    }

    vector position(int i) const { return vector(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

    // particle i from the host mirror (after fetch())
    element particle(int i) const
    {
        element e;
        auto v3 = [i](const std::vector<double> &a) { return vector(a[3 * i], a[3 * i + 1], a[3 * i + 2]); };
        e.r = radius[i];
        e.p = v3(p);
        e.p_old = v3(p_old);
        e.u = v3(u);
        e.u_old = v3(u_old);
        e.f_pp = v3(f_pp);
        e.f_pi = v3(f_pi);
        e.c = c_loc[i];
        e.dcdx = v3(dcdx);
        return e;
    }

    // cluster::distribute, cluster.cpp:53-92, as seeded random sequential
    // addition with a minimum surface gap.  Throws if the box jams.
    void distribute(uint64_t seed = 2, double gap = 0.0)
    {
        const fipi_params &q = config::get();
        const int64_t got = fipi_init_rsa(N, q.r_p, gap, q.lx, q.ly, q.lz, seed, 100000, p.data());
        if (got < 0) check((int)got);
        if (got != N) throw error(FIPI_ERR_ARG, "random sequential addition jammed before placing all spheres");
        p_old = p;  // cluster.cpp:88
        upload();
    }

    // per-particle contact angles (cos theta = 0 is the reference's neutral wetting)
    void set_contact_angles_deg(double lo, double hi, uint64_t seed = 3)
    {
        std::vector<double> th(N);
        check(fipi_init_uniform(N, lo, hi, seed, th.data()));
        const double pi = 3.14159265358979323846;
        for (int i = 0; i < N; i++) cos_theta[i] = std::cos(th[i] * pi / 180.0);
        upload();
    }

    // cluster::sorting, cluster.cpp:116-169: the contact cell list on the
    // device (ids ascending in each cell, as the reference's appends) and
    // it_mark, the first particle of largest y above 0 (kept otherwise)
    void sorting()
    {
        if (!N) return;
        check(fipi_cells_build(device::ctx(), dp.d(), N, static_cast<int32_t *>((void *)dstart.d()),
                               static_cast<int32_t *>((void *)dids.d()), &ncell),
              device::ctx());
        cells_built = true;
        dp.download(p.data(), sizeof(double) * 3 * (size_t)N);
        double y_max = 0.0;
        for (int it = 0; it < N; it++)
            if (p[3 * it + 1] > y_max) {
                y_max = p[3 * it + 1];
                it_mark = it;
            }
    }

    // the reference's cell lists (list of particle ids per contact cell,
    // cluster.hpp:35) from the last sorting()
    std::vector<list> cells() const
    {
        if (!cells_built) throw error(FIPI_ERR_STATE, "cells() before sorting()");
        std::vector<int32_t> start((size_t)ncell + 1), ids(N);
        dstart.download(start.data(), sizeof(int32_t) * start.size());
        dids.download(ids.data(), sizeof(int32_t) * ids.size());
        std::vector<list> out(ncell);
        for (int c = 0; c < ncell; c++)
            for (int k = start[c]; k < start[c + 1]; k++) out[c].append(ids[k]);
        return out;
    }

    // cluster.cpp:307-323: c and grad c at the particle centres (uf is zero,
    // element.cpp:38-52)
    void interpolate_local_variables(const field &u_grid, const field &c, const field &grad_c)
    {
        (void)u_grid;
        const fipi_params &q = config::get();
        if (!N || !q.switch_ch) return;
        fipi_ctx *ctx = device::ctx();
        const size_t G = (size_t)q.nx * q.ny * q.nz;
        check(fipi_interp(ctx, grid_of(c), dp.d(), N, dc.d()), ctx);
        const double *g = grid_of(grad_c);
        std::vector<double> comp(N);
        for (int d = 0; d < 3; d++) {
            check(fipi_interp(ctx, g + d * G, dp.d(), N, dfpi.d()), ctx);  // staging; f_pi is rewritten next
            dfpi.download(comp.data(), sizeof(double) * N);
            for (int i = 0; i < N; i++) dcdx[3 * i + d] = comp[i];
        }
        ddcdx.upload(dcdx.data(), sizeof(double) * 3 * (size_t)N);
        dc.download(c_loc.data(), sizeof(double) * N);
    }

    // cluster.cpp:286-290 / element.cpp:75-97
    void compute_f_pi()
    {
        if (!N) return;
        check(fipi_capillary_forces(device::ctx(), dc.d(), ddcdx.d(), drad.d(), dcth.d(), N, dfpi.d()), device::ctx());
    }

    // cluster.cpp:174-275: f_pp on the device; tau (the packing energy, only
    // reported) on the host from the contact pairs, as the reference sums it
    // over pairs it < id whose f_pi and the marked particle's f_pi are non-zero
    void compute_f_pp(int t)
    {
        if (!N) return;
        fipi_ctx *ctx = device::ctx();
        check(fipi_contact_forces(ctx, dp.d(), N, t, dfpp.d()), ctx);
        int64_t count = 0;
        check(fipi_contact_pairs(ctx, dp.d(), N, nullptr, 0, &count), ctx);
        tau = 0.0;
        if (!count) return;
        dbuf dpairs(sizeof(int32_t) * 2 * (size_t)count);
        check(fipi_contact_pairs(ctx, dp.d(), N, static_cast<int32_t *>((void *)dpairs.d()), count, &count), ctx);
        std::vector<int32_t> pr(2 * (size_t)count);
        dpairs.download(pr.data(), sizeof(int32_t) * pr.size());
        dp.download(p.data(), sizeof(double) * 3 * (size_t)N);
        dfpi.download(f_pi.data(), sizeof(double) * 3 * (size_t)N);
        const fipi_params &q = config::get();
        const double L[3] = {q.lx, q.ly, q.lz}, D[3] = {dx_cell * (nx_cell - 2), dy_cell * (ny_cell - 2),
                                                       dz_cell * (nz_cell - 2)};
        const double rmp = t < 0 ? 0.0 : (t < q.ramp_t ? t / q.ramp_t : 1.0);  // ramp, primitive.cpp:122-129
        auto fnorm = [&](int i) { return position_norm(f_pi, i); };
        const bool mark = fnorm(it_mark) != 0.0;
        for (int64_t k = 0; k < count; k++) {
            const int it = pr[2 * k], id = pr[2 * k + 1];
            double d[3];
            for (int a = 0; a < 3; a++) {
                d[a] = p[3 * it + a] - p[3 * id + a];
                if (d[a] > D[a]) d[a] -= L[a];
                else if (d[a] < -D[a]) d[a] += L[a];
            }
            const double dn = vector(d[0], d[1], d[2]).norm();
            if (!(dn < q.r_eq) || dn == 0.0 || !mark || fnorm(it) == 0.0) continue;
            const double s = (q.r_eq - dn) * rmp * q.k_n;
            double e = 0.0;
            for (int a = 0; a < 3; a++) e += 0.5 * (s * (d[a] / dn)) * d[a];
            tau += e;
        }
    }
    static double position_norm(const std::vector<double> &a, int i)
    {
        return vector(a[3 * i], a[3 * i + 1], a[3 * i + 2]).norm();
    }

    // cluster.cpp:279-283: f_ext is zero (element.cpp:100-103)
    void compute_f_ext(int) {}

    // cluster.cpp:326-341 / element.cpp:156-186
    void advect(const std::string &type)
    {
        if (!N) return;
        int stage;
        if (type == "1") stage = 1;
        else if (type == "2") stage = 2;
        else return;  // the reference does nothing for another type
        check(fipi_advect(device::ctx(), stage, N, dp.d(), dpo.d(), du.d(), duo.d(), dfpi.d(), dfpp.d(), drad.d()),
              device::ctx());
    }
    void enforce_bc()
    {
        if (!N) return;
        check(fipi_advect(device::ctx(), 0, N, dp.d(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr),
              device::ctx());
    }

    // cluster.cpp:344-410: (f_pp + f_ext) / (dx dy dz) spread onto the grid
    field compute_f_r()
    {
        field f_r("VECTORS");
        if (!N) return f_r;
        check(fipi_spread(device::ctx(), dp.d(), dfpp.d(), N, grid.d()), device::ctx());
        f_r.assign_device_real(grid.d());
        return f_r;
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
    // (from the host mirror)
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

    // initial state: phi from `c` (real space), particles from `s`
    // (positions, radii, contact angles of its host mirror); velocity zero
    void load(const field &c, const cluster &s)
    {
        if (c.kspace()) throw error(FIPI_ERR_ARG, "load() takes the phase field in real space");
        const std::vector<double> phi = c.real(0);
        check(fipi_set_state(ctx, phi.data(), s.N ? s.p.data() : nullptr, s.N ? s.radius.data() : nullptr,
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

    // copy the state into facade fields (real space) and a cluster (its
    // host mirror and device buffers); any may be null
    void fetch(field *c, field *u = nullptr, field *grad_c = nullptr, cluster *s = nullptr)
    {
        const fipi_params &q = config::get();
        const size_t G = (size_t)q.nx * q.ny * q.nz;
        std::vector<double> h;
        if (c) {
            h.resize(G);
            check(fipi_get_phi(ctx, h.data()), ctx);
            c->assign_real(h.data(), 0);
        }
        for (int w = 0; w < 2; w++) {
            field *f = w == 0 ? u : grad_c;
            if (!f) continue;
            h.resize(3 * G);
            check(w == 0 ? fipi_get_velocity(ctx, h.data()) : fipi_get_grad_phi(ctx, h.data()), ctx);
            for (int d = 0; d < 3; d++) f->assign_real(h.data() + d * G, d);
        }
        if (s && s->N) {
            check(fipi_get_particles(ctx, s->p.data(), s->p_old.data(), s->u.data(), s->u_old.data(), s->f_pp.data(),
                                     s->f_pi.data()),
                  ctx);
            fipi_diag d{};
            check(fipi_diagnostics(ctx, &d), ctx);
            s->tau = d.tau;
            s->it_mark = d.it_mark;
            s->upload();
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
