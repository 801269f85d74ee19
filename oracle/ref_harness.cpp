// Harness around the reference's own routine headers (test infrastructure).
//
// Compiled by oracle/Makefile with -I pointing at the reference's header
// directory; the headers are used unmodified where they lie.  Each wrapper
// copies plain arrays into the std::vector arguments the reference routines
// take, calls the routine, and copies the result back.  Only the headers that
// need nothing beyond the C++ standard library are included (commutators.h,
// init_G.h, AS.h and RK_solver.h need MKL headers this image lacks).

#include <iostream>
#include <complex>
#include <cstring>
#include <string>
#include <vector>
#include <unistd.h>

#include "HF.h"
#include "collision_integral.h"
#include "psi.h"
#include "pi.h"
#include "G1_G2_t_step.h"
#include "fermi_func.h"
#include "dipole.h"
#include "read_input.h"
#include "write_to_file.h"

typedef std::complex<double> cd;

static std::vector<cd> vin(const double* p, long n) {
    std::vector<cd> v(n);
    if (p) std::memcpy(v.data(), p, sizeof(cd) * n);
    return v;
}
static void vout(double* p, const std::vector<cd>& v) { std::memcpy(p, v.data(), sizeof(cd) * v.size()); }

extern "C" {

void ref_hf(double* h, const double* G1, int Ns, int Nb, double U, double t1, double t2, double w,
            const double* eps, const double* V, int nV, int quench_extent) {
    const long n2 = (long)Ns * Nb * Ns * Nb;
    std::vector<cd> hv = vin(h, n2), gv = vin(G1, n2);
    std::vector<double> ev(eps, eps + Nb), Vv(V, V + nV);
    h_HFx_quench(hv, gv, gv, Ns, Nb, U, t1, t2, w, ev, Vv, quench_extent);
    vout(h, hv);
}

void ref_collision(double* I, const double* P, const double* Q, int Ns, int Nb, double U,
                   const double* V, int nV) {
    const long n2 = (long)Ns * Nb * Ns * Nb, n4 = n2 * n2;
    std::vector<cd> iv(n2), pv = vin(P, n4), qv = vin(Q, n4);
    std::vector<double> Vv(V, V + nV);
    collision_int(iv, pv, qv, Ns, Nb, U, Vv);
    vout(I, iv);
}

void ref_psi(double* psi, const double* G1, int Ns, int Nb, double U, const double* V, int nV,
             int same_spin) {
    const long n2 = (long)Ns * Nb * Ns * Nb, n4 = n2 * n2;
    std::vector<cd> sv(n4), gv = vin(G1, n2);
    std::vector<double> Vv(V, V + nV);
    if (same_spin)
        Psi_xxxx(sv, gv, Ns, Nb, U, Vv);
    else
        Psi_xyxy(sv, gv, Ns, Nb, U, Vv);
    vout(psi, sv);
}

void ref_pi(double* pi, const double* G1, const double* P, const double* Q, double U, int Ns, int Nb,
            const double* V, int nV) {
    const long n2 = (long)Ns * Nb * Ns * Nb, n4 = n2 * n2;
    std::vector<cd> ov(n4), gv = vin(G1, n2), pv = vin(P, n4), qv = vin(Q, n4);
    std::vector<double> Vv(V, V + nV);
    Pi_xyxy(ov, gv, pv, qv, U, Ns, Nb, Vv);
    vout(pi, ov);
}

void ref_g1_step(const double* comm, const double* I, double* K, double* RK, double w, int Ns, int Nb) {
    const long n2 = (long)Ns * Nb * Ns * Nb;
    std::vector<cd> cv = vin(comm, n2), iv = vin(I, n2), kv = vin(K, n2), rv = vin(RK, n2);
    G1_step(cv, iv, kv, rv, w, Ns, Nb);
    vout(K, kv);
    vout(RK, rv);
}

void ref_g2_step(const double* comm, const double* psi, const double* pi, double* K, double* RK,
                 double w, int Ns, int Nb) {
    const long n2 = (long)Ns * Nb * Ns * Nb, n4 = n2 * n2;
    std::vector<cd> cv = vin(comm, n4), sv = vin(psi, n4), pv = vin(pi, n4), kv = vin(K, n4),
                    rv = vin(RK, n4);
    G2_step(cv, sv, pv, kv, rv, w, Ns, Nb);
    vout(K, kv);
    vout(RK, rv);
}

double ref_fermi(double t, double t0, double tau) { return fermi(t, t0, tau); }

void ref_dipole(double* p, const double* traj, long T, int N) {
    std::vector<std::vector<cd>> G(T);
    for (long t = 0; t < T; ++t) G[t] = vin(traj + t * 2L * N * N, (long)N * N);
    std::vector<double> pv(T);
    dipole(pv, G);
    std::memcpy(p, pv.data(), sizeof(double) * T);
}

// Parsed values of the reference's input globals.
struct ref_inputs {
    int HF, EHM;
    double t1, t2, t3, U;
    int time_steps;
    double dt_fixed, quench_strength;
    int Ns, Nq, Nb;
    int q_type;  // 0 full, 1 pulse, 2 other
    double AS_rate, AS_midpoint, quench_rate, quench_on_midpoint, quench_off_midpoint;
};

// Runs the reference parser on <dir>/input (it reads ./input).
int ref_parse(const char* dir, ref_inputs* out) {
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return -1;
    if (chdir(dir) != 0) return -2;
    // restore the compiled defaults first; the parser only assigns
    HF = false; EHM = false; t1 = 1; t2 = 0; t3 = 0; U = .5; time_steps = 1000; dt_fixed = .02;
    quench_strength = 1; Ns = 2; Nq = 0; Nb = 1; q_type = "full"; AS_rate = 3; AS_midpoint = 25;
    quench_rate = .2; quench_on_midpoint = 50; quench_off_midpoint = 55;
    int rc = 0;
    try {
        assign_vals();
    } catch (...) {
        rc = -3;
    }
    if (chdir(cwd) != 0) return -4;
    out->HF = HF; out->EHM = EHM; out->t1 = t1; out->t2 = t2; out->t3 = t3; out->U = U;
    out->time_steps = time_steps; out->dt_fixed = dt_fixed; out->quench_strength = quench_strength;
    out->Ns = Ns; out->Nq = Nq; out->Nb = Nb;
    out->q_type = q_type == "full" ? 0 : (q_type == "pulse" ? 1 : 2);
    out->AS_rate = AS_rate; out->AS_midpoint = AS_midpoint; out->quench_rate = quench_rate;
    out->quench_on_midpoint = quench_on_midpoint; out->quench_off_midpoint = quench_off_midpoint;
    return rc;
}

// Writes real_G1.txt / imag_G1.txt and dipole.txt in <dir> with the reference writers.
int ref_write(const char* dir, const double* traj, long T, int N, const double* p) {
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return -1;
    if (chdir(dir) != 0) return -2;
    std::vector<std::vector<cd>> G(T);
    for (long t = 0; t < T; ++t) G[t] = vin(traj + t * 2L * N * N, (long)N * N);
    write_to_cfile2D("G1.txt", G);
    write_to_rfile("dipole.txt", std::vector<double>(p, p + T));
    return chdir(cwd) == 0 ? 0 : -4;
}

}  // extern "C"
