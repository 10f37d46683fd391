// Driver in the shape of the reference's main() (src/main.cpp:19-252) with
// the time-loop body (main.cpp:75-222) replaced by the GPU step.
//
//   fipi_main [n] [n_particles] [t_end] [precision 0|1] [write_vtk 0|1]
//
// Defaults: 64^3 box in benchmark units, 500 spheres of radius 2.5 dx,
// 200 steps, FP64, no files.  With write_vtk=1 the phase field and the
// particles are written to data/ every t_end/100 steps (main.cpp:224-241).
#include <chrono>
#include <cstdlib>
#include <iostream>

#include "fipi.hpp"

int main(int argc, char **argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 64;
    const int np = argc > 2 ? std::atoi(argv[2]) : 500;
    const int t_end = argc > 3 ? std::atoi(argv[3]) : 200;
    const int prec = argc > 4 ? std::atoi(argv[4]) : FIPI_FP64;
    const bool vtk = argc > 5 && std::atoi(argv[5]) != 0;
    try {
        fipi::config::set(fipi::benchmark_params(n, np, 2.5, t_end, prec));

        // main.cpp:34 (phi0 = U(-0.05, 0.05) from a seed instead of rand())
        fipi::field c("SCALARS", "phase_field", 1, 0.0, 0.05);
        fipi::field u("VECTORS");
        fipi::field grad_c("VECTORS");
        if (vtk) c.output_to_file(fipi::convert_to_vtk(0, "phase"), "phase");

        // main.cpp:58-66
        fipi::cluster suspension;
        suspension.summary();
        if (suspension.N) {
            suspension.distribute(2, 0.5);
            if (vtk) suspension.output(fipi::convert_to_vtk(0, "particle"));
            std::cout << "\n";
        }

        fipi::solver gpu;
        gpu.load(c, suspension);
        const auto start = std::chrono::steady_clock::now();
        for (int t = 0; t < t_end; t++) {
            gpu.step(1);  // main.cpp:75-222
            if (fipi::output_time(t, t_end)) {
                const fipi_diag d = gpu.diagnostics();
                std::cout << "Time = " << (t + 1) * fipi::config::get().dt << " ";
                std::cout << "total free energy = " << d.free_energy << " ";
                if (vtk) {
                    gpu.fetch(&c, &u, &grad_c, &suspension);
                    c.output_to_file(fipi::convert_to_vtk(t + 1, "phase"), "phase");
                    if (suspension.N) suspension.output(fipi::convert_to_vtk(t + 1, "particle"));
                } else if (suspension.N) {
                    std::cout << "packing energy= " << d.tau << " ";
                }
                std::cout << "\n";
            }
        }
        const double s =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        gpu.fetch(&c, nullptr, nullptr, &suspension);
        const fipi_diag d = gpu.diagnostics();
        std::cout.precision(17);
        std::cout << "steps " << gpu.time() << " wall " << s << " s, " << t_end / s << " steps/s, mean phi "
                  << d.mean_phi << " free energy " << d.free_energy << "\n";
    } catch (const fipi::error &e) {
        std::cerr << e.what() << " (status " << e.code << ")\n";
        return 1;
    }
    return 0;
}
