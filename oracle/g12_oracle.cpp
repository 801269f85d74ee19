// CPU oracle for the HF-GKBA G1-G2 (GW) time propagation.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain C++ restatement of the
// reference's routines, written element by element from the equations the
// reference code evaluates, in the reference's own memory layout and with the
// reference's loop complexity (the O(N^5) two-particle commutator, O(N^4 * N)
// Psi/Pi, OpenMP on the same routines the reference parallelises).  It is the
// checker for the HIP library and the CPU baseline timed beside it; nothing in
// the product may link or call it.
//
// Layouts (reference, G12KBA-GW/src/headers/commutators.h:59-60):
//   one-particle index  X = site + orb*Ns, N = Ns*Nb
//   G1[X][Y]            at X*N + Y
//   pair index          A(p1,p2) = i + j*Ns + (a + b*Nb)*Ns*Ns
//                       for p1 = (i,a), p2 = (j,b)
//   G2[A][B]            at A*N*N + B       (A = (p1,p2) row, B = (q1,q2) column)
// Complex numbers are interleaved (re, im) doubles, i.e. std::complex<double>.
//
// Interaction tables used below (the reference spells them out as loops):
//   W_xy(X,Z) = U         if site(X) == site(Z)
//             = V[n-1]    if |site(X)-site(Z)| = n, 1 <= n <= nV
//             = 0         otherwise
//   W_xx(X,Z) = W_xy(X,Z) except W_xx(X,X) = 0.
//
// Every routine cites the reference function it restates.

#include <complex>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

typedef std::complex<double> cplx;
static const cplx I1(0.0, 1.0);

namespace {

struct Geo {
    int Ns, Nb, N;
    std::vector<int64_t> pr;  // reference pair index of (p1, p2) at p1*N + p2
    Geo(int ns, int nb) : Ns(ns), Nb(nb), N(ns * nb), pr((size_t)ns * nb * ns * nb) {
        for (int p1 = 0; p1 < N; ++p1)
            for (int p2 = 0; p2 < N; ++p2) {
                const int i = p1 % Ns, a = p1 / Ns, j = p2 % Ns, b = p2 / Ns;
                pr[(size_t)p1 * N + p2] =
                    (int64_t)i + (int64_t)j * Ns + ((int64_t)a + (int64_t)b * Nb) * Ns * Ns;
            }
    }
    int site(int x) const { return x % Ns; }
    int orb(int x) const { return x / Ns; }
    int one(int s, int o) const { return s + o * Ns; }
    int64_t pair(int p1, int p2) const { return pr[(size_t)p1 * N + p2]; }
    int64_t g2(int p1, int p2, int q1, int q2) const {
        return pair(p1, p2) * (int64_t)N * N + pair(q1, q2);
    }
};

inline double w_xy(const Geo& G, double U, const double* V, int nV, int x, int z) {
    int d = std::abs(G.site(x) - G.site(z));
    if (d == 0) return U;
    if (d <= nV) return V[d - 1];
    return 0.0;
}
inline double w_ss(const Geo& G, double U, const double* V, int nV, int x, int z, bool same_spin) {
    if (same_spin && x == z) return 0.0;
    return w_xy(G, U, V, nV, x, z);
}

inline const cplx* C(const double* p) { return reinterpret_cast<const cplx*>(p); }
inline cplx* C(double* p) { return reinterpret_cast<cplx*>(p); }

}  // namespace

extern "C" {

// Hartree-Fock one-body Hamiltonian, accumulated into h (caller zeroes it).
// Restates h_HFx_quench, G12KBA-GW/src/headers/HF.h:57-104: orbital energy and
// quench shift on the diagonal of sites < quench_extent, -t1 between nearest
// sites of the same orbital, -t2 between adjacent orbitals of one site, onsite
// Hartree -iU g_XX - 2iU sum_{g!=a} g_(i,g)(i,g), onsite inter-orbital Fock
// +iU g[(i,g)][(i,a)] into h[(i,a)][(i,g)] (transposed element, HF.h:83),
// long-range Hartree -2iV_n sum_d g over both neighbours at distance n and
// long-range Fock +iV_n g[P][Q] into h[P][Q] (HF.h:91-93).
void orc_hf(double* h_, const double* G1_, int Ns, int Nb, double U, double t1, double t2,
            double w, const double* eps, const double* V, int nV, int quench_extent) {
    Geo G(Ns, Nb);
    const int N = G.N;
    cplx* h = C(h_);
    const cplx* g = C(G1_);
    auto H = [&](int r, int c) -> cplx& { return h[(int64_t)r * N + c]; };
    auto g1 = [&](int r, int c) { return g[(int64_t)r * N + c]; };
    for (int i = 0; i < Ns; ++i) {
        for (int a = 0; a < Nb; ++a) {
            const int X = G.one(i, a);
            H(X, X) += (eps ? eps[a] : 0.0);
            if (i < quench_extent) H(X, X) += w;
            if (i + 1 < Ns) {
                H(G.one(i + 1, a), X) -= t1;
                H(X, G.one(i + 1, a)) -= t1;
            }
            if (a + 1 < Nb) {
                H(X, G.one(i, a + 1)) -= t2;
                H(G.one(i, a + 1), X) -= t2;
            }
            H(X, X) += -I1 * U * g1(X, X);
            for (int o = 0; o < Nb; ++o) {
                if (o == a) continue;
                const int Z = G.one(i, o);
                H(X, Z) += I1 * U * g1(Z, X);
                H(X, X) += -2.0 * I1 * U * g1(Z, Z);
            }
            for (int n = 1; n <= nV; ++n) {
                const double v = V[n - 1];
                for (int side = -1; side <= 1; side += 2) {
                    const int s2 = i + side * n;
                    if (s2 < 0 || s2 >= Ns) continue;
                    for (int o = 0; o < Nb; ++o) {
                        const int Z = G.one(s2, o);
                        H(X, X) += -2.0 * I1 * v * g1(Z, Z);
                        // Fock between X and every orbital of the site at distance n
                        // (each ordered pair once over the i loop).
                        if (side > 0) {
                            H(G.one(s2, a), G.one(i, o)) += I1 * v * g1(G.one(s2, a), G.one(i, o));
                            H(X, Z) += I1 * v * g1(X, Z);
                        }
                    }
                }
            }
        }
    }
}

// [h, G] for N x N matrices.  Restates hG_commutator,
// G12KBA-GW/src/headers/commutators.h:28-51 (two zgemm then subtraction).
void orc_comm1(const double* G_, const double* h_, double* comm_, int N) {
    const cplx* g = C(G_);
    const cplx* h = C(h_);
    cplx* out = C(comm_);
    for (int r = 0; r < N; ++r)
        for (int c = 0; c < N; ++c) {
            cplx hg = 0.0, gh = 0.0;
            for (int z = 0; z < N; ++z) {
                hg += h[(int64_t)r * N + z] * g[(int64_t)z * N + c];
                gh += g[(int64_t)r * N + z] * h[(int64_t)z * N + c];
            }
            out[(int64_t)r * N + c] = hg - gh;
        }
}

// Two-particle commutator [hx (x) 1 + 1 (x) hx, M] with the right-hand
// Hamiltonian hy, element by element:
//   C[A;B] = sum_Z hx(p2,Z) M[(p1,Z);B] + hx(p1,Z) M[(Z,p2);B]
//          - M[A;(q1,Z)] hy(Z,q2) - M[A;(Z,q2)] hy(Z,q1).
// Restates h2xyG2xyxy_commutator, G12KBA-GW/src/headers/commutators.h:53-99
// (4 N^5 complex multiply-adds, OpenMP over the output as the reference).
void orc_comm2(const double* M_, const double* hx_, const double* hy_, double* C_, int Ns, int Nb) {
    Geo G(Ns, Nb);
    const int N = G.N;
    const cplx* M = C(M_);
    const cplx* hx = C(hx_);
    const cplx* hy = C(hy_);
    cplx* out = C(C_);
#pragma omp parallel for collapse(2) schedule(static)
    for (int p1 = 0; p1 < N; ++p1)
        for (int p2 = 0; p2 < N; ++p2)
            for (int q1 = 0; q1 < N; ++q1)
                for (int q2 = 0; q2 < N; ++q2) {
                    cplx acc = 0.0;
                    for (int z = 0; z < N; ++z) {
                        acc += hx[(int64_t)p2 * N + z] * M[G.g2(p1, z, q1, q2)];
                        acc += hx[(int64_t)p1 * N + z] * M[G.g2(z, p2, q1, q2)];
                        acc -= M[G.g2(p1, p2, q1, z)] * hy[(int64_t)z * N + q2];
                        acc -= M[G.g2(p1, p2, z, q2)] * hy[(int64_t)z * N + q1];
                    }
                    out[G.g2(p1, p2, q1, q2)] = acc;
                }
}

// Collision integral I[X,Y] = -i sum_Z ( W_xy(X,Z) P[(X,Z);(Y,Z)] + W_xx(X,Z) Q[(X,Z);(Y,Z)] )
// with P the first G2 argument (xyxy) and Q the second (xxxx).
// Restates collision_int, G12KBA-GW/src/headers/collision_integral.h:27-69.
void orc_collision(double* I_, const double* P_, const double* Q_, int Ns, int Nb, double U,
                   const double* V, int nV) {
    Geo G(Ns, Nb);
    const int N = G.N;
    const cplx* P = C(P_);
    const cplx* Q = C(Q_);
    cplx* I = C(I_);
    for (int X = 0; X < N; ++X)
        for (int Y = 0; Y < N; ++Y) {
            cplx acc = 0.0;
            for (int Z = 0; Z < N; ++Z) {
                const double wxy = w_ss(G, U, V, nV, X, Z, false);
                const double wxx = w_ss(G, U, V, nV, X, Z, true);
                if (wxy == 0.0 && wxx == 0.0) continue;
                const int64_t e = G.g2(X, Z, Y, Z);
                acc += wxy * P[e] + wxx * Q[e];
            }
            I[(int64_t)X * N + Y] = -I1 * acc;
        }
}

// GW two-particle source, element by element (same_spin=1: Psi_xxxx with
// W_xx, same_spin=0: Psi_xyxy with W_xy):
//   Psi[A;B] = g(p1,q1) g(p2,q2) (W(p1,p2) - W(q1,q2))
//            + i g(p1,q1) sum_Z (W(p1,Z) - W(q1,Z)) g(p2,Z) g(Z,q2)
//            + i g(p2,q2) sum_Z (W(p2,Z) - W(q2,Z)) g(p1,Z) g(Z,q1)
// Restates Psi_xxxx / Psi_xyxy, G12KBA-GW/src/headers/psi.h:157-277 / 26-154
// (the onsite orbital exclusion r != a of psi.h:194-195 is W_xx(X,X) = 0; the
// extra xyxy lines psi.h:63-72 restore the r == a terms, i.e. W_xy(X,X) = U).
void orc_psi(double* psi_, const double* G1_, int Ns, int Nb, double U, const double* V, int nV,
             int same_spin) {
    Geo G(Ns, Nb);
    const int N = G.N;
    const cplx* g = C(G1_);
    cplx* psi = C(psi_);
    const bool ss = same_spin != 0;
    auto g1 = [&](int r, int c) { return g[(int64_t)r * N + c]; };
    // sites that carry a non-zero interaction with site s: |d| <= nV
#pragma omp parallel for collapse(2) schedule(static)
    for (int p1 = 0; p1 < N; ++p1)
        for (int p2 = 0; p2 < N; ++p2)
            for (int q1 = 0; q1 < N; ++q1)
                for (int q2 = 0; q2 < N; ++q2) {
                    const cplx g11 = g1(p1, q1), g22 = g1(p2, q2);
                    cplx v = g11 * g22 * (w_ss(G, U, V, nV, p1, p2, ss) - w_ss(G, U, V, nV, q1, q2, ss));
                    cplx s1 = 0.0, s2 = 0.0;
                    for (int Z = 0; Z < N; ++Z) {
                        const double a1 = w_ss(G, U, V, nV, p1, Z, ss) - w_ss(G, U, V, nV, q1, Z, ss);
                        if (a1 != 0.0) s1 += a1 * g1(p2, Z) * g1(Z, q2);
                        const double a2 = w_ss(G, U, V, nV, p2, Z, ss) - w_ss(G, U, V, nV, q2, Z, ss);
                        if (a2 != 0.0) s2 += a2 * g1(p1, Z) * g1(Z, q1);
                    }
                    v += I1 * g11 * s1 + I1 * g22 * s2;
                    psi[G.g2(p1, p2, q1, q2)] = v;
                }
}

// GW polarisation term Pi[A;B] = g(p2,q2) (F[p1,q1;p2] - F[p1,q1;q2]) with
//   F[X,Y;Z] = -i sum_Z' ( W_xy(Z,Z') P[(X,Z');(Y,Z')] + W_xx(Z,Z') Q[(X,Z');(Y,Z')] ),
// P the first G2 argument, Q the second.  Restates Pi_xyxy,
// G12KBA-GW/src/headers/pi.h:25-95 (called with swapped G2 pairs for the two
// spin channels, G12KBA-GW/src/headers/RK_solver.h:163-164).
void orc_pi(double* pi_, const double* G1_, const double* P_, const double* Q_, double U, int Ns,
            int Nb, const double* V, int nV) {
    Geo G(Ns, Nb);
    const int N = G.N;
    const cplx* g = C(G1_);
    const cplx* P = C(P_);
    const cplx* Q = C(Q_);
    cplx* pi = C(pi_);
#pragma omp parallel for collapse(2) schedule(static)
    for (int p1 = 0; p1 < N; ++p1)
        for (int p2 = 0; p2 < N; ++p2)
            for (int q1 = 0; q1 < N; ++q1)
                for (int q2 = 0; q2 < N; ++q2) {
                    cplx f = 0.0;
                    for (int Z = 0; Z < N; ++Z) {
                        const double xa = w_ss(G, U, V, nV, p2, Z, false) - w_ss(G, U, V, nV, q2, Z, false);
                        const double xb = w_ss(G, U, V, nV, p2, Z, true) - w_ss(G, U, V, nV, q2, Z, true);
                        if (xa == 0.0 && xb == 0.0) continue;
                        const int64_t e = G.g2(p1, Z, q1, Z);
                        f += xa * P[e] + xb * Q[e];
                    }
                    pi[G.g2(p1, p2, q1, q2)] = g[(int64_t)p2 * N + q2] * (-I1 * f);
                }
}

// K = -i (comm + I + I^dagger);  RK += w K.
// Restates G1_step, G12KBA-GW/src/headers/G1_G2_t_step.h:24-47.
void orc_g1_step(const double* comm_, const double* I_, double* K_, double* RK_, double w, int Ns,
                 int Nb) {
    const int N = Ns * Nb;
    const cplx* cm = C(comm_);
    const cplx* I = C(I_);
    cplx* K = C(K_);
    cplx* RK = C(RK_);
    for (int X = 0; X < N; ++X)
        for (int Y = 0; Y < N; ++Y) {
            const int64_t e = (int64_t)X * N + Y;
            K[e] = -I1 * (cm[e] + I[e] + std::conj(I[(int64_t)Y * N + X]));
            RK[e] += w * K[e];
        }
}

// K[A;B] = -i (comm + psi + pi - conj(pi[B^x;A^x]));  RK += w K,
// where A^x swaps the two particles of a pair.
// Restates G2_step, G12KBA-GW/src/headers/G1_G2_t_step.h:55-86.
void orc_g2_step(const double* comm_, const double* psi_, const double* pi_, double* K_,
                 double* RK_, double w, int Ns, int Nb) {
    Geo G(Ns, Nb);
    const int N = G.N;
    const cplx* cm = C(comm_);
    const cplx* ps = C(psi_);
    const cplx* pi = C(pi_);
    cplx* K = C(K_);
    cplx* RK = C(RK_);
    for (int p1 = 0; p1 < N; ++p1)
        for (int p2 = 0; p2 < N; ++p2)
            for (int q1 = 0; q1 < N; ++q1)
                for (int q2 = 0; q2 < N; ++q2) {
                    const int64_t e = G.g2(p1, p2, q1, q2);
                    const int64_t t = G.g2(q2, q1, p2, p1);
                    K[e] = -I1 * (cm[e] + ps[e] + pi[e] - std::conj(pi[t]));
                    RK[e] += w * K[e];
                }
}

// Logistic ramp 1/(1+exp((t0-t)/tau)).  Restates fermi,
// G12KBA-GW/src/headers/fermi_func.h:21-24.
double orc_fermi(double t, double t0, double tau) { return 1.0 / (1.0 + std::exp((t0 - t) / tau)); }

// Dipole trajectory p_t = sum_{i<N/2} 2 (N/2 - i - 1/2)(Im G_t[i,i] - Im G_t[N-1-i,N-1-i]) / N,
// accumulated over i in the outer loop as the reference.
// Restates dipole, G12KBA-GW/src/headers/dipole.h:25-33.
void orc_dipole(double* p, const double* traj, int64_t T, int N) {
    const double n = (double)N;
    for (int64_t t = 0; t < T; ++t) p[t] = 0.0;
    for (int i = 0; i < N / 2; ++i)
        for (int64_t t = 0; t < T; ++t) {
            const cplx* G = C(traj) + t * (int64_t)N * N;
            const double lo = G[(int64_t)i * (N + 1)].imag();
            const double hi = G[(int64_t)N * N - (int64_t)i * (N + 1) - 1].imag();
            p[t] += 2 * (n / 2 - i - .5) * (lo - hi) / n;
        }
}

}  // extern "C"

// ----------------------------------------------------------------------------
// Drivers: adiabatic switching and production RK4 with the reference's loop
// semantics (G12KBA-GW/src/headers/AS.h:63-214, RK_solver.h:64-226).

struct orc_params {
    int Ns, Nb;
    double t1, t2, U;
    double dt;
    int time_steps;
    double quench_strength;
    int Nq;
    int q_pulse;   // 1: pulse protocol, 0: full
    int hf_only;   // HF=true: no correlation (G2 frozen, I = 0)
    double AS_rate, AS_midpoint;
    double quench_rate, quench_on, quench_off;
    double AS_dt;  // 0.1 in the reference
    int AS_tail;   // 20 in the reference
    double gate_time;  // 50 in the reference
    int nV;            // number of long-range shells (0: onsite only)
    const double* V_profile;  // V_n = U_eff * V_profile[n-1]
    const double* eps;        // Nb orbital energies or NULL (zeros)
};

namespace {

struct Work {
    int N;
    int64_t n2, n4;
    std::vector<cplx> h, c1, I, Y1, K1, A1;
    std::vector<cplx> Yxx, Yxy, Kxx, Kxy, Axx, Axy, cxx, cxy, sxx, sxy, pxx, pxy;
    std::vector<double> V;
    explicit Work(int Ns, int Nb) {
        N = Ns * Nb;
        n2 = (int64_t)N * N;
        n4 = n2 * n2;
        for (auto* v : {&h, &c1, &I, &Y1, &K1, &A1}) v->assign(n2, 0.0);
        for (auto* v : {&Yxx, &Yxy, &Kxx, &Kxy, &Axx, &Axy, &cxx, &cxy, &sxx, &sxy, &pxx, &pxy})
            v->assign(n4, 0.0);
    }
};

double* D(std::vector<cplx>& v) { return reinterpret_cast<double*>(v.data()); }

// One RK4 step of the coupled system.  stage_U(k) and stage_q(k, &w, &nq)
// give the interaction strength and the quench for stage k.
template <class FU, class FQ>
void rk4_step(const orc_params& p, Work& W, cplx* G1, cplx* Gxx, cplx* Gxy, double dt_full,
              FU stage_U, FQ stage_q, int nstages = 4) {
    const int Ns = p.Ns, Nb = p.Nb;
    const bool corr = (p.U != 0.0) && !p.hf_only;
    std::fill(W.A1.begin(), W.A1.end(), 0.0);
    std::fill(W.K1.begin(), W.K1.end(), 0.0);
    std::fill(W.Axx.begin(), W.Axx.end(), 0.0);
    std::fill(W.Axy.begin(), W.Axy.end(), 0.0);
    std::fill(W.Kxx.begin(), W.Kxx.end(), 0.0);
    std::fill(W.Kxy.begin(), W.Kxy.end(), 0.0);
    double dtk = dt_full;
    for (int k = 0; k < nstages; ++k) {
        dtk = (k == 1 || k == 2) ? dt_full / 2 : dt_full;
        const double wk = (k == 1 || k == 2) ? 2.0 : 1.0;
        for (int64_t e = 0; e < W.n2; ++e) W.Y1[e] = G1[e] + W.K1[e] * dtk;
        for (int64_t e = 0; e < W.n4; ++e) {
            W.Yxx[e] = Gxx[e] + W.Kxx[e] * dtk;
            W.Yxy[e] = Gxy[e] + W.Kxy[e] * dtk;
        }
        const double Ue = stage_U(k, dtk);
        W.V.assign(p.nV, 0.0);
        for (int n = 0; n < p.nV; ++n) W.V[n] = Ue * p.V_profile[n];
        const double* V = W.V.data();
        double wq = 0.0;
        int nq = 0;
        stage_q(k, dtk, &wq, &nq);
        std::fill(W.h.begin(), W.h.end(), 0.0);
        orc_hf(D(W.h), D(W.Y1), Ns, Nb, Ue, p.t1, p.t2, wq, p.eps, V, p.nV, nq);
        orc_comm1(D(W.Y1), D(W.h), D(W.c1), W.N);
        if (corr) {
            orc_collision(D(W.I), D(W.Yxy), D(W.Yxx), Ns, Nb, Ue, V, p.nV);
            orc_comm2(D(W.Yxx), D(W.h), D(W.h), D(W.cxx), Ns, Nb);
            orc_comm2(D(W.Yxy), D(W.h), D(W.h), D(W.cxy), Ns, Nb);
            orc_psi(D(W.sxx), D(W.Y1), Ns, Nb, Ue, V, p.nV, 1);
            orc_psi(D(W.sxy), D(W.Y1), Ns, Nb, Ue, V, p.nV, 0);
            orc_pi(D(W.pxx), D(W.Y1), D(W.Yxy), D(W.Yxx), Ue, Ns, Nb, V, p.nV);
            orc_pi(D(W.pxy), D(W.Y1), D(W.Yxx), D(W.Yxy), Ue, Ns, Nb, V, p.nV);
            orc_g2_step(D(W.cxx), D(W.sxx), D(W.pxx), D(W.Kxx), D(W.Axx), wk, Ns, Nb);
            orc_g2_step(D(W.cxy), D(W.sxy), D(W.pxy), D(W.Kxy), D(W.Axy), wk, Ns, Nb);
        }
        orc_g1_step(D(W.c1), D(W.I), D(W.K1), D(W.A1), wk, Ns, Nb);
    }
    // final update with the last stage's step (= dt_full), (A*dt)/6
    for (int64_t e = 0; e < W.n2; ++e) G1[e] += W.A1[e] * dtk / 6.0;
    for (int64_t e = 0; e < W.n4; ++e) {
        Gxx[e] += W.Axx[e] * dtk / 6.0;
        Gxy[e] += W.Axy[e] * dtk / 6.0;
    }
}

}  // namespace

extern "C" {

// Adiabatic switching: RK4 at fixed step AS_dt with U_AS(t) = fermi(t; mid, rate) U
// (long-range shells scaled with U_AS), no quench, until AS_tail iterations have
// begun with U_AS == U exactly.  The very first check sees no previous U_AS and
// does not count (the reference reads an uninitialised value there).
// Restates Adiabatic_switching, G12KBA-GW/src/headers/AS.h:63-214.
int orc_adiabatic_switching(const orc_params* pp, double* G1_, double* Gxx_, double* Gxy_,
                            int* steps_out) {
    const orc_params& p = *pp;
    Work W(p.Ns, p.Nb);
    int count = 0, steps = 0;
    bool have = false;
    double last_U = 0.0, t = 0.0;
    while (count < p.AS_tail) {
        if (have && std::fabs(last_U - p.U) <= 0) count += 1;
        auto sU = [&](int k, double dtk) {
            const double ts = (k == 0) ? t : t + dtk;
            last_U = orc_fermi(ts, p.AS_midpoint, p.AS_rate) * p.U;
            return last_U;
        };
        auto sQ = [&](int, double, double* w, int* nq) { *w = 0.0; *nq = 0; };
        rk4_step(p, W, C(G1_), C(Gxx_), C(Gxy_), p.AS_dt, sU, sQ);
        have = true;
        t += p.AS_dt;
        ++steps;
    }
    if (steps_out) *steps_out = steps;
    return 0;
}

// Production RK4: while (t < time_steps*dt) record G1, check the density guard
// (|2 Im G1[0,0]| > 3 or NaN returns 1 after recording), take one step with the
// quench gated off for stage times below gate_time, then t += dt.
// traj receives up to traj_cap recorded G1 matrices; *steps_out the number of
// iterations begun.  Restates RK4, G12KBA-GW/src/headers/RK_solver.h:64-226.
int orc_rk4(const orc_params* pp, double* G1_, double* Gxx_, double* Gxy_, double* traj,
            int64_t traj_cap, int64_t* steps_out) {
    const orc_params& p = *pp;
    Work W(p.Ns, p.Nb);
    const double tmax = double(p.time_steps) * p.dt;
    double t = 0.0;
    int64_t n = 0;
    int rc = 0;
    cplx* G1 = C(G1_);
    while (t < tmax) {
        if (traj && n < traj_cap)
            std::memcpy(traj + n * W.n2 * 2, G1_, sizeof(double) * 2 * W.n2);
        ++n;
        const double dens = 2 * G1[0].imag();
        if (std::fabs(dens) > 3 || std::isnan(dens)) {
            rc = 1;
            break;
        }
        auto sU = [&](int, double) { return p.U; };
        auto sQ = [&](int k, double dtk, double* w, int* nq) {
            const double ts = (k == 0) ? t : t + dtk;
            if (ts < p.gate_time) {
                *w = 0.0;
                *nq = 0;
                return;
            }
            const double on = orc_fermi(ts, p.quench_on, p.quench_rate);
            if (p.q_pulse)
                *w = (on * (1 - orc_fermi(ts, p.quench_off, p.quench_rate))) * p.quench_strength;
            else
                *w = on * p.quench_strength;
            *nq = p.Nq;
        };
        rk4_step(p, W, G1, C(Gxx_), C(Gxy_), p.dt, sU, sQ);
        t += p.dt;
    }
    if (steps_out) *steps_out = n;
    return rc;
}

// One RK4 step at fixed stage scalars (no time dependence): for unit tests of
// the stepping scheme against closed forms.
void orc_rk4_fixed_step(const orc_params* pp, double* G1_, double* Gxx_, double* Gxy_, double dt) {
    const orc_params& p = *pp;
    Work W(p.Ns, p.Nb);
    auto sU = [&](int, double) { return p.U; };
    auto sQ = [&](int, double, double* w, int* nq) { *w = 0.0; *nq = 0; };
    rk4_step(p, W, C(G1_), C(Gxx_), C(Gxy_), dt, sU, sQ);
}

// The first nstages stages of one RK4 step (stage input, derivative, RK
// accumulation; the state update only after stage 4).  Used to time a
// bounded sample of the CPU path.
void orc_rk4_partial_step(const orc_params* pp, double* G1_, double* Gxx_, double* Gxy_, double dt,
                          int nstages) {
    const orc_params& p = *pp;
    Work W(p.Ns, p.Nb);
    auto sU = [&](int, double) { return p.U; };
    auto sQ = [&](int, double, double* w, int* nq) { *w = 0.0; *nq = 0; };
    rk4_step(p, W, C(G1_), C(Gxx_), C(Gxy_), dt, sU, sQ, nstages);
}

}  // extern "C"
