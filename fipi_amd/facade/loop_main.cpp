// The reference's time-loop body (src/main.cpp:75-222) written statement by
// statement through the facade's classes -- every field operator, transform
// and cluster method runs in libfipi -- next to the fused GPU step
// (fipi::solver, fipi_step) from the same initial state.  After `steps`
// steps it prints the relative L2 differences of phi, u and the particle
// positions and the packing energies of the two, and exits 0 when
// phi, u and positions agree to 1e-12 and tau to 1e-10, 2 otherwise.
//
//   fipi_loop [n] [n_particles] [steps] [t_end]
//
// Defaults: 32^3 in benchmark units, 300 spheres of radius 2.5 dx placed
// with a surface gap of -1 dx (so that they touch), contact angles
// U(60, 120) deg, phi0 = U(-0.05, 0.05), 5 steps of a run of t_end = 20
// (the contact force ramps over the first 4 steps).
#include <chrono>
#include <cstdlib>
#include <iostream>

#include "fipi.hpp"

using fipi::cluster;
using fipi::field;

namespace {

double rel_l2(const std::vector<double> &a, const std::vector<double> &b)
{
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < a.size(); i++) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

std::vector<double> components(const field &f)
{
    std::vector<double> out;
    for (int d = 0; d < f.dimension(); d++) {
        const std::vector<double> r = f.real(d);
        out.insert(out.end(), r.begin(), r.end());
    }
    return out;
}

}  // namespace

int main(int argc, char **argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 32;
    const int np = argc > 2 ? std::atoi(argv[2]) : 300;
    const int steps = argc > 3 ? std::atoi(argv[3]) : 5;
    const int t_end = argc > 4 ? std::atoi(argv[4]) : 20;
    try {
        fipi::config::set(fipi::benchmark_params(n, np, 2.5, t_end));
        const fipi_params &q = fipi::config::get();
        const double lambda = q.lambda, epsilon = q.epsilon, dt = q.dt, mobility = q.mobility, mu = q.mu;
        const double rho1 = q.rho1, rho2 = q.rho2;
        const bool switch_ch = q.switch_ch, switch_ns = q.switch_ns, switch_particle = q.switch_particle;
        auto sqr = [](double a) { return a * a; };

        // main.cpp:28-66
        field kf("VECTORS", "wave_vector");
        field kf2("SCALARS");
        kf2 = kf.power(2.0);

        field c("SCALARS", "phase_field", 1, 0.0, 0.05);
        field c_new("SCALARS");
        field c_old("SCALARS");
        field f("SCALARS");
        field f_old("SCALARS");
        field grad_c("VECTORS");
        field conv_c("SCALARS");
        field conv_c_old("SCALARS");

        field u("VECTORS");
        field pch("SCALARS");
        field g("VECTORS", "gravity");
        field f_total("VECTORS");

        cluster suspension;
        if (switch_particle) {
            suspension.distribute(2, -1.0);
            suspension.set_contact_angles_deg(60.0, 120.0, 3);
        }

        // the fused step from the same state
        fipi::solver fused;
        fused.load(c, suspension);

        const fipi::plan p_forward = fipi::plan_forward();
        const fipi::plan p_backward = fipi::plan_backward();

        const auto start = std::chrono::steady_clock::now();
        for (int t = 0; t < steps; t++) {
            // **************************** Phase Field (main.cpp:79-148)
            if (switch_ch) {
                f = (lambda / sqr(epsilon)) * (c.power(3.0) - c);

                f.fft(p_forward);
                c.fft(p_forward);

                if (switch_ns) {
                    grad_c = c.grad(kf);
                    grad_c.ifft(p_backward);

                    conv_c = u.dot(grad_c);
                    conv_c.fft(p_forward);
                }

                if (t == 0) {
                    c_old = c;
                    f_old = f;

                    if (switch_ns) conv_c_old = conv_c;

                    c -= dt * mobility * kf2 * f;

                    if (switch_ns) c -= dt * conv_c;

                    c /= 1.0 + (dt * mobility * lambda) * (kf2.power(2.0));
                } else {
                    c_new = 4.0 * c - c_old;
                    c_new -= (4.0 * dt * mobility) * kf2 * f;
                    c_new += (2.0 * dt * mobility) * kf2 * f_old;

                    if (switch_ns) {
                        c_new -= (4.0 * dt) * conv_c;
                        c_new += (2.0 * dt) * conv_c_old;
                    }

                    c_new /= 3.0 + (2.0 * dt * mobility * lambda) * (kf2.power(2.0));

                    c_old = c;
                    f_old = f;

                    if (switch_ns) conv_c_old = conv_c;

                    c = c_new;
                }

                c.ifft(p_backward);
            }

            // ****************************** Fluid Velocity Field (main.cpp:153-190)
            if (switch_ns) {
                f_total.reset("p");

                if (switch_ch) {
                    f = (lambda / sqr(epsilon)) * (c.power(3.0) - c);
                    f.fft(p_forward);
                    c.fft(p_forward);

                    pch = lambda * kf2 * c + f;
                    pch.ifft(p_backward);

                    grad_c = c.grad(kf);
                    grad_c.ifft(p_backward);
                    c.ifft(p_backward);

                    f_total += pch * grad_c;
                }

                if (switch_particle) f_total += suspension.compute_f_r();

                f_total += ((1.0 + c) / 2.0 * rho1 + (1.0 - c) / 2.0 * rho2) * g;

                f_total.fft(p_forward);

                u = f_total - kf * (kf.dot(f_total)) / kf2;
                u /= mu * kf2;
                u.ifft(p_backward);
            }

            // ********************** Particles (main.cpp:195-222)
            if (switch_particle) {
                if (switch_ch) {
                    c.fft(p_forward);
                    grad_c = c.grad(kf);
                    grad_c.ifft(p_backward);
                    c.ifft(p_backward);
                }

                suspension.interpolate_local_variables(u, c, grad_c);
                suspension.sorting();

                suspension.compute_f_pi();
                suspension.compute_f_pp(t + 1);
                suspension.compute_f_ext(t + 1);

                suspension.advect("1");
                suspension.enforce_bc();

                suspension.interpolate_local_variables(u, c, grad_c);
                suspension.sorting();

                suspension.compute_f_pi();
                suspension.compute_f_pp(t + 1);
                suspension.compute_f_ext(t + 1);

                suspension.advect("2");
                suspension.enforce_bc();
            }
        }
        const double loop_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();

        fused.step(steps);
        field c_f("SCALARS"), u_f("VECTORS");
        cluster s_f;
        fused.fetch(&c_f, &u_f, nullptr, &s_f);
        suspension.fetch();

        const double e_phi = rel_l2(c.real(0), c_f.real(0));
        const double e_u = rel_l2(components(u), components(u_f));
        double e_pos = 0.0, e_tau = 0.0;
        if (suspension.N) {
            // minimum-image differences of the positions
            const double L[3] = {q.lx, q.ly, q.lz};
            std::vector<double> d(suspension.p.size()), z(suspension.p.size(), 0.0);
            for (size_t i = 0; i < d.size(); i++) {
                const double a = suspension.p[i] - s_f.p[i];
                d[i] = a - L[i % 3] * std::nearbyint(a / L[i % 3]);
            }
            e_pos = rel_l2(d, z) / std::max(rel_l2(s_f.p, z), 1e-300);
            e_tau = s_f.tau != 0.0 ? std::abs(suspension.tau / s_f.tau - 1.0) : std::abs(suspension.tau);
        }
        std::cout.precision(6);
        std::cout << std::scientific;
        std::cout << "steps " << steps << " facade loop " << loop_s << " s\n";
        std::cout << "phi " << e_phi << " u " << e_u << " positions " << e_pos << " tau " << suspension.tau << " vs "
                  << s_f.tau << " (" << e_tau << ") it_mark " << suspension.it_mark << " vs " << s_f.it_mark << "\n";
        const bool ok = e_phi < 1e-12 && e_u < 1e-12 && e_pos < 1e-12 && e_tau < 1e-10 &&
                        suspension.it_mark == s_f.it_mark;
        std::cout << (ok ? "MATCH" : "MISMATCH") << "\n";
        return ok ? 0 : 2;
    } catch (const fipi::error &e) {
        std::cerr << e.what() << " (status " << e.code << ")\n";
        return 1;
    }
}
