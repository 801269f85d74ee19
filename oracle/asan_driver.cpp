// Test infrastructure: drives the host-side C++ of the boundary
// (g12kba_amd/csrc/host_io.cpp: input parser, writers, G0, checkpoints,
// dipole) and the CPU oracle's drivers (g12_oracle.cpp, g12_fast.cpp) under
// AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile, target
// `sanitize`).  The reference itself reads uninitialised values
// (RK_solver.h:80 t_temp, AS.h:77 U_AS); the restatements must not.
//
//   asan_driver DIR   parses DIR/input_* (one case per file), writes outputs
//                     and a checkpoint into DIR, runs the oracle drivers;
//                     prints "parsed <name> <values...>" per case and exits 0
//                     when every self-check holds.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <dirent.h>
#include <random>
#include <string>
#include <vector>

#include "../g12kba_amd/csrc/host_io.h"
#include "../include/g12kba.h"

typedef std::complex<double> cplx;

// oracle entry points (g12_oracle.cpp, g12_fast.cpp)
struct orc_params {
    int Ns, Nb;
    double t1, t2, U;
    double dt;
    int time_steps;
    double quench_strength;
    int Nq;
    int q_pulse;
    int hf_only;
    double AS_rate, AS_midpoint;
    double quench_rate, quench_on, quench_off;
    double AS_dt;
    int AS_tail;
    double gate_time;
    int nV;
    const double* V_profile;
    const double* eps;
};
extern "C" int orc_adiabatic_switching(const orc_params* pp, double* G1, double* Gxx, double* Gxy, int* steps);
extern "C" int orc_rk4(const orc_params* pp, double* G1, double* Gxx, double* Gxy, double* traj, int64_t traj_cap,
                       int64_t* steps_out);
extern "C" int orcf_run(const void* pp, double* G1, double* Gxx, double* Gxy, int run_as, int* as_steps,
                        int64_t* rk_steps, double* traj, int64_t traj_cap, int traj_every, double* dipole_src,
                        int64_t dens_cap);

static int failures = 0;
#define CHECK(cond)                                                        \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
            ++failures;                                                    \
        }                                                                  \
    } while (0)

static double maxdiff(const std::vector<double>& a, const std::vector<double>& b) {
    double m = 0;
    for (size_t i = 0; i < a.size(); ++i) m = std::fmax(m, std::fabs(a[i] - b[i]));
    return m;
}

static void parse_cases(const std::string& dir) {
    DIR* d = opendir(dir.c_str());
    CHECK(d != nullptr);
    if (!d) return;
    std::vector<std::string> names;
    while (dirent* e = readdir(d))
        if (std::strncmp(e->d_name, "input_", 6) == 0) names.push_back(e->d_name);
    closedir(d);
    CHECK(!names.empty());
    for (const std::string& n : names) {
        g12_params p;
        g12_params_default(&p);
        const int rc = g12_parse_input((dir + "/" + n).c_str(), &p);
        std::printf("parsed %s %d %d %d %.17g %.17g %.17g %.17g %d %.17g %.17g %d %d %d %d %.17g %.17g %.17g %.17g %.17g\n",
                    n.c_str() + 6, rc, p.HF, p.EHM, p.t1, p.t2, p.t3, p.U, p.time_steps, p.dt, p.quench_strength,
                    p.Ns, p.Nq, p.Nb, p.q_type, p.AS_rate, p.AS_midpoint, p.quench_rate, p.quench_on_midpoint,
                    p.quench_off_midpoint);
    }
    g12_params p;
    g12_params_default(&p);
    CHECK(g12_parse_input((dir + "/no_such_input").c_str(), &p) == 0);  // a missing file gives zeros
}

static void host_io_round_trips(const std::string& dir) {
    g12_params p;
    g12_params_default(&p);
    p.Ns = 5;
    p.Nb = 2;
    p.t2 = 0.4;
    const int N = p.Ns * p.Nb;
    const int64_t n2 = (int64_t)N * N, n4 = n2 * n2;
    std::vector<double> G1(2 * n2), G1f(2 * n2), Gw(2 * n2);
    CHECK(g12_make_g0(&p, G1.data(), 0) == 0);
    CHECK(g12_make_g0(&p, G1f.data(), 1) == 0);
    CHECK(g12_init_g_test(&p, Gw.data()) == 0);
    double tr = 0;
    for (int i = 0; i < N; ++i) tr += G1[2 * (i * N + i) + 1];
    CHECK(std::fabs(tr - p.Ns / 2) < 1e-12);  // filling Ns/2 (Main.cpp:52)
    // a short trajectory: text outputs and dipole
    const int T = 7;
    std::vector<double> traj(2 * n2 * T), dip(T);
    for (int s = 0; s < T; ++s)
        for (int64_t e = 0; e < 2 * n2; ++e) traj[s * 2 * n2 + e] = G1[e] * (1.0 + 0.01 * s);
    CHECK(g12_dipole(dip.data(), traj.data(), T, N) == 0);
    CHECK(g12_write_outputs(dir.c_str(), traj.data(), T, N, dip.data(), 42.5) == 0);
    // checkpoint round trip
    std::mt19937_64 rng(7);
    std::normal_distribution<double> nd;
    std::vector<double> Gxx(2 * n4), Gxy(2 * n4), a(2 * n2), bxx(2 * n4), bxy(2 * n4);
    for (auto& x : Gxx) x = nd(rng);
    for (auto& x : Gxy) x = nd(rng);
    const std::string ck = dir + "/state.bin";
    CHECK(g12io::save_checkpoint(ck.c_str(), p.Ns, p.Nb, 1.25, 125, G1.data(), Gxx.data(), Gxy.data()) == 0);
    double t = 0;
    int64_t steps = 0;
    CHECK(g12io::load_checkpoint(ck.c_str(), p.Ns, p.Nb, &t, &steps, a.data(), bxx.data(), bxy.data()) == 0);
    CHECK(t == 1.25 && steps == 125 && a == G1 && bxx == Gxx && bxy == Gxy);
    CHECK(g12io::load_checkpoint(ck.c_str(), p.Ns + 1, p.Nb, &t, &steps, a.data(), bxx.data(), bxy.data()) != 0);
    // eigen-decomposition of a random Hermitian matrix: A v = v w
    const int n = 9;
    std::vector<double> A(2 * n * n), A0, w(n), v(2 * n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            const cplx z(nd(rng), i == j ? 0.0 : nd(rng));
            A[2 * (i * n + j)] = z.real();
            A[2 * (i * n + j) + 1] = z.imag();
            A[2 * (j * n + i)] = z.real();
            A[2 * (j * n + i) + 1] = -z.imag();
        }
    A0 = A;
    CHECK(g12io::hermitian_eigen(A.data(), n, w.data(), v.data()) == 0);
    const cplx* M = reinterpret_cast<const cplx*>(A0.data());
    const cplx* V = reinterpret_cast<const cplx*>(v.data());
    double res = 0;
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < n; ++i) {
            cplx s = 0;
            for (int j = 0; j < n; ++j) s += M[i * n + j] * V[j * n + k];
            res = std::fmax(res, std::abs(s - w[k] * V[i * n + k]));
        }
    CHECK(res < 1e-10);
}

static void oracle_drivers() {
    const int Ns = 3, Nb = 2, N = Ns * Nb;
    const int64_t n2 = (int64_t)N * N, n4 = n2 * n2;
    const double prof[2] = {0.5, 0.25};
    orc_params p{};
    p.Ns = Ns;
    p.Nb = Nb;
    p.t1 = 1.0;
    p.t2 = 0.4;
    p.U = 1.2;
    p.dt = 0.02;
    p.time_steps = 12;
    p.quench_strength = 0.8;
    p.Nq = 1;
    p.q_pulse = 1;
    p.AS_rate = 0.05;
    p.AS_midpoint = 0.1;
    p.quench_rate = 0.05;
    p.quench_on = 0.1;
    p.quench_off = 0.2;
    p.AS_dt = 0.1;
    p.AS_tail = 3;
    p.gate_time = 0.05;
    p.nV = 2;
    p.V_profile = prof;
    std::vector<double> G1(2 * n2, 0.0), Gxx(2 * n4, 0.0), Gxy(2 * n4, 0.0);
    for (int i = 0; i < N / 2; ++i) G1[2 * (i * N + i) + 1] = 1.0;
    std::vector<double> fG1 = G1, fGxx = Gxx, fGxy = Gxy;
    int nas = 0;
    CHECK(orc_adiabatic_switching(&p, G1.data(), Gxx.data(), Gxy.data(), &nas) == 0);
    std::vector<double> traj(2 * n2 * (p.time_steps + 2));
    int64_t nrk = 0;
    CHECK(orc_rk4(&p, G1.data(), Gxx.data(), Gxy.data(), traj.data(), p.time_steps + 2, &nrk) == 0);
    int fas = 0;
    int64_t frk = 0;
    CHECK(orcf_run(&p, fG1.data(), fGxx.data(), fGxy.data(), 1, &fas, &frk, nullptr, 0, 0, nullptr, 0) == 0);
    int64_t want = 0;  // iterations of while (t < tmax) with accumulated t (RK_solver.h:96,216)
    for (double t = 0.0; t < double(p.time_steps) * p.dt; t += p.dt) ++want;
    CHECK(nas == fas && nrk == frk && nrk == want);
    CHECK(maxdiff(G1, fG1) < 1e-13 && maxdiff(Gxx, fGxx) < 1e-13 && maxdiff(Gxy, fGxy) < 1e-13);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: asan_driver DIR\n");
        return 2;
    }
    const std::string dir = argv[1];
    parse_cases(dir);
    host_io_round_trips(dir);
    oracle_drivers();
    std::printf("checks failed: %d\n", failures);
    return failures ? 1 : 0;
}
