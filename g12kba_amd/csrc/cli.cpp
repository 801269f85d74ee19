// g12kba — command-line drop-in for the reference program (Main.cpp:44-62):
// parse ./input, build G1(0), adiabatic switching, RK4 with the G1 trajectory,
// dipole, and write real_G1.txt / imag_G1.txt / dipole.txt / progress.txt.
//
//   g12kba [input-file] [--g0-fp32] [--g0-test] [--save-state FILE] [--load-state FILE] [--g0 FILE]
//          [--steps N]
//
// --g0 FILE reads G1(0) (N*N complex128, raw binary) instead of the
// eigen-solver; --g0-test starts from the domain-wall state of init_G_test
// (first half of the chain filled); --load-state skips G0 and adiabatic
// switching and continues the production loop from the checkpoint's time and
// step count; --steps N stops after at most N production steps (with
// --save-state, a run can be split and resumed bit for bit).  The outputs
// hold the steps run by this invocation.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/g12kba.h"

int main(int argc, char** argv) {
    std::string input = "input", save, load, g0file;
    int fp32 = 0, g0test = 0;
    long long max_steps = -1;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--g0-fp32")
            fp32 = 1;
        else if (a == "--g0-test")
            g0test = 1;
        else if (a == "--save-state" && i + 1 < argc)
            save = argv[++i];
        else if (a == "--load-state" && i + 1 < argc)
            load = argv[++i];
        else if (a == "--g0" && i + 1 < argc)
            g0file = argv[++i];
        else if (a == "--steps" && i + 1 < argc)
            max_steps = std::atoll(argv[++i]);
        else
            input = a;
    }
    g12_params p;
    g12_params_default(&p);
    if (g12_parse_input(input.c_str(), &p) != 0) {
        std::fprintf(stderr, "g12kba: could not parse %s\n", input.c_str());
        return 2;
    }
    const int N = p.Ns * p.Nb;
    const int64_t n2 = (int64_t)N * N;
    std::vector<double> G1(2 * n2, 0.0);
    g12_ctx* ctx = nullptr;
    int rc = g12_create(&p, &ctx);
    if (rc) {
        std::fprintf(stderr, "g12kba: context creation failed (%d)\n", rc);
        return 3;
    }
    if (!load.empty()) {
        if ((rc = g12_load_state(ctx, load.c_str()))) {
            std::fprintf(stderr, "g12kba: cannot load %s (%d)\n", load.c_str(), rc);
            return 4;
        }
    } else {
        if (!g0file.empty()) {
            FILE* f = std::fopen(g0file.c_str(), "rb");
            if (!f || std::fread(G1.data(), sizeof(double) * 2, n2, f) != (size_t)n2) {
                std::fprintf(stderr, "g12kba: cannot read %s\n", g0file.c_str());
                return 4;
            }
            std::fclose(f);
        } else if (g0test) {
            g12_init_g_test(&p, G1.data());
        } else {
            g12_make_g0(&p, G1.data(), fp32);
        }
        g12_set_state(ctx, G1.data(), nullptr, nullptr);
        std::cout << "Beginning Adiabatic Switching Procedure\n";
        auto t0 = std::chrono::steady_clock::now();
        int as_steps = 0;
        if ((rc = g12_adiabatic_switching(ctx, &as_steps))) return 5;
        auto t1 = std::chrono::steady_clock::now();
        std::cout << "Adiabatic Switching complete, elapsed time:" << std::chrono::duration<double>(t1 - t0).count()
                  << " sec \n";
    }
    const double tmax = double(p.time_steps) * p.dt;
    // the production loop from the context's clock (0 after switching, the
    // checkpoint's after --load-state): iterations of while (t < tmax) with
    // t accumulated as in RK_solver.h:96,216
    double t_start = 0.0;
    g12_get_time(ctx, &t_start, nullptr);
    int64_t cap = 0;
    for (double t = t_start; t < tmax; t += p.dt) ++cap;
    const bool partial = max_steps >= 0 && max_steps < cap;
    if (partial) cap = max_steps;
    std::vector<double> traj((size_t)(cap > 0 ? cap : 1) * 2 * n2);
    std::cout << "Beginning Time Stepping Procedure\n";
    auto t0 = std::chrono::steady_clock::now();
    int64_t steps = 0;
    if (load.empty() && !partial)
        rc = g12_rk4(ctx, traj.data(), cap, &steps);  // the reference's protocol from t = 0
    else
        rc = g12_rk4_record(ctx, cap, traj.data(), cap, &steps);
    auto t1 = std::chrono::steady_clock::now();
    if (rc == G12_GUARD) {
        // the density of the failing record (RK_solver.h:100-110)
        const double dens = steps > 0 ? 2 * traj[(size_t)(steps - 1) * 2 * n2 + 1] : 0.0;
        std::cout << "ERROR: Unphysical Number Density on site 1:" << dens
                  << " elapsed time:" << std::chrono::duration<double>(t1 - t0).count() << " sec \n";
    } else if (rc < 0) {
        std::fprintf(stderr, "g12kba: RK4 failed (%d)\n", rc);
        return 6;
    } else {
        std::cout << "\n Execution Succesful, elapsed time:" << std::chrono::duration<double>(t1 - t0).count()
                  << " sec \n";
    }
    std::vector<double> dip(steps > 0 ? steps : 1);
    g12_dipole(dip.data(), traj.data(), steps, N);
    // progress.txt holds t / tmax * 100 after the last completed step
    // (RK_solver.h:216-220); none is written when no step completed
    double t_end = 0;
    g12_get_time(ctx, &t_end, nullptr);
    const int64_t completed = rc == G12_GUARD ? steps - 1 : steps;
    const double progress = completed > 0 ? t_end / tmax * 100 : -1.0;
    if (g12_write_outputs(".", traj.data(), steps, N, dip.data(), progress)) return 7;
    if (!save.empty()) g12_save_state(ctx, save.c_str());
    g12_destroy(ctx);
    return 0;
}
