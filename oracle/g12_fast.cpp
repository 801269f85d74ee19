// Factorised FP64 CPU oracle (test infrastructure only).
//
// Same mathematics as g12_oracle.cpp, evaluated through the partial-trace
// tables so that long runs at N = 32 finish on a CPU (the element-wise oracle
// needs ~2 s per RK4 step there).  G2 tensors are in the natural layout
// T[p1][p2][q1][q2] (pair index p1*N + p2, p = site + orb*Ns); the drivers
// take and return reference-layout arrays and permute once.
//
//   [H2,M]      = hx on p1 + hx on p2 - hy on q1 (transposed) - hy on q2
//                 (commutators.h:53-99)
//   D^s[X,Y,Z]  = T^s[X][Z][Y][Z]
//   F^xx        = -i (D^xy Wxy^T + D^xx Wxx^T),  F^xy = -i (D^xx Wxy^T + D^xy Wxx^T)
//                 (pi.h:25-95 with the argument swap of RK_solver.h:163-164)
//   I[X,Y]      = F^xx[X,Y,X]                     (collision_integral.h:27-69)
//   Q^s[Z,X,Y]  = sum_w g[X,w] W^s[Z,w] g[w,Y]
//   Psi, Pi, Pi-dagger and K as in G1_G2_t_step.h:55-86 / psi.h:26-277.
// Its results agree with g12_oracle.cpp to rounding (checked by the CPU suite).

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

typedef std::complex<double> cplx;

namespace {

struct orc_params_f {  // identical layout to orc_params in g12_oracle.cpp
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

}  // namespace

extern "C" void orc_hf(double* h_, const double* G1_, int Ns, int Nb, double U, double t1, double t2, double w,
                       const double* eps, const double* V, int nV, int quench_extent);
extern "C" double orc_fermi(double t, double t0, double tau);

namespace {

struct Fast {
    int Ns, Nb, N;
    int64_t n2, n3, n4;
    std::vector<int64_t> pr;  // natural (p1*N+p2) -> reference pair index
    std::vector<double> Wxy, Wxx;
    std::vector<cplx> h, Y1, K1, A1, c1, I, D[2], F[2], Q[2];
    std::vector<cplx> Y[2], K[2], A[2];
    Fast(int ns, int nb) : Ns(ns), Nb(nb), N(ns * nb) {
        n2 = (int64_t)N * N;
        n3 = n2 * N;
        n4 = n2 * n2;
        pr.resize(n2);
        for (int p1 = 0; p1 < N; ++p1)
            for (int p2 = 0; p2 < N; ++p2) {
                const int i = p1 % Ns, a = p1 / Ns, j = p2 % Ns, b = p2 / Ns;
                pr[(size_t)p1 * N + p2] = i + (int64_t)j * Ns + ((int64_t)a + (int64_t)b * Nb) * Ns * Ns;
            }
        Wxy.assign(n2, 0.0);
        Wxx.assign(n2, 0.0);
        for (auto* v : {&h, &Y1, &K1, &A1, &c1, &I}) v->assign(n2, 0.0);
        for (int s = 0; s < 2; ++s) {
            D[s].assign(n3, 0.0);
            F[s].assign(n3, 0.0);
            Q[s].assign(n3, 0.0);
            Y[s].assign(n4, 0.0);
            K[s].assign(n4, 0.0);
            A[s].assign(n4, 0.0);
        }
    }
    void to_nat(const cplx* ref, cplx* nat) const {
#pragma omp parallel for schedule(static)
        for (int64_t A_ = 0; A_ < n2; ++A_)
            for (int64_t B = 0; B < n2; ++B) nat[A_ * n2 + B] = ref[pr[A_] * n2 + pr[B]];
    }
    void to_ref(const cplx* nat, cplx* ref) const {
#pragma omp parallel for schedule(static)
        for (int64_t A_ = 0; A_ < n2; ++A_)
            for (int64_t B = 0; B < n2; ++B) ref[pr[A_] * n2 + pr[B]] = nat[A_ * n2 + B];
    }
    void w_tables(double U, const double* V, int nV) {
        for (int x = 0; x < N; ++x)
            for (int z = 0; z < N; ++z) {
                const int d = std::abs(x % Ns - z % Ns);
                const double w = d == 0 ? U : (d <= nV ? V[d - 1] : 0.0);
                Wxy[(size_t)x * N + z] = w;
                Wxx[(size_t)x * N + z] = x == z ? 0.0 : w;
            }
    }
    // out = [H2, T] with hx = hy = h (natural layout)
    void comm2(const cplx* T, cplx* out) const {
        const cplx* H = h.data();
#pragma omp parallel for schedule(static)
        for (int p1 = 0; p1 < N; ++p1) {
            cplx* o = out + (int64_t)p1 * n3;
            std::memset(static_cast<void*>(o), 0, sizeof(cplx) * n3);
            // h on p2
            for (int p2 = 0; p2 < N; ++p2)
                for (int z = 0; z < N; ++z) {
                    const cplx hz = H[(int64_t)p2 * N + z];
                    const cplx* t = T + ((int64_t)p1 * N + z) * n2;
                    cplx* oo = o + (int64_t)p2 * n2;
                    for (int64_t c = 0; c < n2; ++c) oo[c] += hz * t[c];
                }
            // h on p1
            for (int z = 0; z < N; ++z) {
                const cplx hz = H[(int64_t)p1 * N + z];
                const cplx* t = T + (int64_t)z * n3;
                for (int64_t c = 0; c < n3; ++c) o[c] += hz * t[c];
            }
            // h on q1 and q2 (from the right)
            for (int p2 = 0; p2 < N; ++p2)
                for (int q1 = 0; q1 < N; ++q1) {
                    cplx* oo = o + ((int64_t)p2 * N + q1) * N;
                    const cplx* trow = T + (((int64_t)p1 * N + p2) * N + q1) * N;
                    for (int z = 0; z < N; ++z) {
                        const cplx tz = trow[z];
                        const cplx hq1 = H[(int64_t)z * N + q1];
                        const cplx* hrow = H + (int64_t)z * N;
                        const cplx* tq = T + (((int64_t)p1 * N + p2) * N + z) * N;
                        for (int q2 = 0; q2 < N; ++q2) oo[q2] -= tz * hrow[q2] + hq1 * tq[q2];
                    }
                }
        }
    }
    void tables(bool corr) {
        if (!corr) return;
        for (int s = 0; s < 2; ++s) {
            const cplx* T = Y[s].data();
            for (int X = 0; X < N; ++X)
                for (int Yy = 0; Yy < N; ++Yy)
                    for (int Z = 0; Z < N; ++Z)
                        D[s][((int64_t)X * N + Yy) * N + Z] = T[(((int64_t)X * N + Z) * N + Yy) * N + Z];
        }
        const cplx mi(0.0, -1.0);
#pragma omp parallel for schedule(static)
        for (int64_t xy = 0; xy < n2; ++xy)
            for (int Z = 0; Z < N; ++Z) {
                cplx fxx = 0.0, fxy = 0.0;
                for (int w = 0; w < N; ++w) {
                    const double a = Wxy[(size_t)Z * N + w], b = Wxx[(size_t)Z * N + w];
                    const cplx dxx = D[0][xy * N + w], dxy = D[1][xy * N + w];
                    fxx += a * dxy + b * dxx;
                    fxy += a * dxx + b * dxy;
                }
                F[0][xy * N + Z] = mi * fxx;
                F[1][xy * N + Z] = mi * fxy;
            }
        for (int X = 0; X < N; ++X)
            for (int Yy = 0; Yy < N; ++Yy) I[(size_t)X * N + Yy] = F[0][((int64_t)X * N + Yy) * N + X];
        const cplx* g = Y1.data();
#pragma omp parallel for collapse(2) schedule(static)
        for (int Z = 0; Z < N; ++Z)
            for (int X = 0; X < N; ++X)
                for (int Yy = 0; Yy < N; ++Yy) {
                    cplx axy = 0.0, axx = 0.0;
                    for (int w = 0; w < N; ++w) {
                        const double wv = Wxy[(size_t)Z * N + w];
                        if (wv == 0.0) continue;
                        const cplx t = g[(size_t)X * N + w] * g[(size_t)w * N + Yy] * wv;
                        axy += t;
                        if (w != Z) axx += t;
                    }
                    Q[0][((int64_t)Z * N + X) * N + Yy] = axx;
                    Q[1][((int64_t)Z * N + X) * N + Yy] = axy;
                }
    }
    // K^s = -i (comm + Psi + Pi - Pi^dagger) into Kout, from comm in Kout
    void rhs(int s, cplx* Kout) const {
        const cplx* g = Y1.data();
        const cplx* Fs = F[s].data();
        const cplx* Qs = Q[s].data();
        const std::vector<double>& W = s == 0 ? Wxx : Wxy;
        const cplx I1(0.0, 1.0);
#pragma omp parallel for collapse(2) schedule(static)
        for (int p1 = 0; p1 < N; ++p1)
            for (int p2 = 0; p2 < N; ++p2)
                for (int q1 = 0; q1 < N; ++q1)
                    for (int q2 = 0; q2 < N; ++q2) {
                        const int64_t e = (((int64_t)p1 * N + p2) * N + q1) * N + q2;
                        const cplx g11 = g[(size_t)p1 * N + q1], g22 = g[(size_t)p2 * N + q2];
                        cplx psi = g11 * g22 * (W[(size_t)p1 * N + p2] - W[(size_t)q1 * N + q2]);
                        psi += I1 * g11 * (Qs[((int64_t)p1 * N + p2) * N + q2] - Qs[((int64_t)q1 * N + p2) * N + q2]);
                        psi += I1 * g22 * (Qs[((int64_t)p2 * N + p1) * N + q1] - Qs[((int64_t)q2 * N + p1) * N + q1]);
                        const cplx pi = g22 * (Fs[((int64_t)p1 * N + q1) * N + p2] - Fs[((int64_t)p1 * N + q1) * N + q2]);
                        const cplx pt = g[(size_t)q1 * N + p1] *
                                        (Fs[((int64_t)q2 * N + p2) * N + q1] - Fs[((int64_t)q2 * N + p2) * N + p1]);
                        Kout[e] = -I1 * (Kout[e] + psi + pi - std::conj(pt));
                    }
    }
};

template <class FU, class FQ>
void step(const orc_params_f& p, Fast& W, cplx* G1, cplx* S[2], double dt_full, FU stage_U, FQ stage_q) {
    const bool corr = p.U != 0.0 && !p.hf_only;
    const int Ns = p.Ns, Nb = p.Nb;
    std::vector<double> V(p.nV);
    double dtk = dt_full;
    std::vector<cplx> Knew(corr ? W.n4 : 1);
    for (int k = 0; k < 4; ++k) {
        dtk = (k == 1 || k == 2) ? dt_full / 2 : dt_full;
        const double wk = (k == 1 || k == 2) ? 2.0 : 1.0;
        for (int64_t e = 0; e < W.n2; ++e) W.Y1[e] = G1[e] + (k ? W.K1[e] : cplx(0.0)) * dtk;
        if (corr)
            for (int s = 0; s < 2; ++s) {
                cplx* Ys = W.Y[s].data();
                const cplx* Ks = W.K[s].data();
                const cplx* Ss = S[s];
#pragma omp parallel for schedule(static)
                for (int64_t e = 0; e < W.n4; ++e) Ys[e] = Ss[e] + (k ? Ks[e] : cplx(0.0)) * dtk;
            }
        const double Ue = stage_U(k, dtk);
        for (int n = 0; n < p.nV; ++n) V[n] = Ue * p.V_profile[n];
        double wq = 0.0;
        int nq = 0;
        stage_q(k, dtk, &wq, &nq);
        std::fill(W.h.begin(), W.h.end(), 0.0);
        orc_hf(reinterpret_cast<double*>(W.h.data()), reinterpret_cast<const double*>(W.Y1.data()), Ns, Nb, Ue, p.t1,
               p.t2, wq, p.eps, V.data(), p.nV, nq);
        W.w_tables(Ue, V.data(), p.nV);
        const int N = W.N;
        for (int X = 0; X < N; ++X)
            for (int Yy = 0; Yy < N; ++Yy) {
                cplx hg = 0.0, gh = 0.0;
                for (int z = 0; z < N; ++z) {
                    hg += W.h[(size_t)X * N + z] * W.Y1[(size_t)z * N + Yy];
                    gh += W.Y1[(size_t)X * N + z] * W.h[(size_t)z * N + Yy];
                }
                W.c1[(size_t)X * N + Yy] = hg - gh;
            }
        W.tables(corr);
        const cplx mi(0.0, -1.0);
        for (int X = 0; X < N; ++X)
            for (int Yy = 0; Yy < N; ++Yy) {
                const size_t e = (size_t)X * N + Yy;
                cplx v = W.c1[e];
                if (corr) v += W.I[e] + std::conj(W.I[(size_t)Yy * N + X]);
                W.K1[e] = mi * v;
                W.A1[e] = (k ? W.A1[e] : cplx(0.0)) + wk * W.K1[e];
            }
        if (corr)
            for (int s = 0; s < 2; ++s) {
                W.comm2(W.Y[s].data(), Knew.data());
                W.rhs(s, Knew.data());
                cplx* As = W.A[s].data();
                const cplx* kn = Knew.data();
#pragma omp parallel for schedule(static)
                for (int64_t e = 0; e < W.n4; ++e) As[e] = (k ? As[e] : cplx(0.0)) + wk * kn[e];
                W.K[s].swap(Knew);
            }
    }
    for (int64_t e = 0; e < W.n2; ++e) G1[e] += W.A1[e] * dtk / 6.0;
    if (corr)
        for (int s = 0; s < 2; ++s) {
            const cplx* As = W.A[s].data();
            cplx* Ss = S[s];
#pragma omp parallel for schedule(static)
            for (int64_t e = 0; e < W.n4; ++e) Ss[e] += As[e] * dtk / 6.0;
        }
}

}  // namespace

extern "C" {

// Adiabatic switching + RK4 with the loop semantics of the element-wise oracle
// (AS.h:63-214, RK_solver.h:64-226).  G2 in the reference layout.  run_as = 0
// skips adiabatic switching.  traj_every > 0 records G1 every traj_every-th
// iteration (plus the first) into traj (capacity traj_cap).  dens_out, if not
// NULL, receives Im G1[0,0] before every iteration (capacity dens_cap).
int orcf_run(const void* pp, double* G1_, double* Gxx_, double* Gxy_, int run_as, int* as_steps,
             int64_t* rk_steps, double* traj, int64_t traj_cap, int traj_every, double* dipole_src, int64_t dens_cap) {
    const orc_params_f& p = *static_cast<const orc_params_f*>(pp);
    Fast W(p.Ns, p.Nb);
    cplx* G1 = reinterpret_cast<cplx*>(G1_);
    std::vector<cplx> Snat[2] = {std::vector<cplx>(W.n4), std::vector<cplx>(W.n4)};
    W.to_nat(reinterpret_cast<cplx*>(Gxx_), Snat[0].data());
    W.to_nat(reinterpret_cast<cplx*>(Gxy_), Snat[1].data());
    cplx* S[2] = {Snat[0].data(), Snat[1].data()};
    if (run_as) {
        int count = 0, n = 0;
        bool have = false;
        double last = 0.0, t = 0.0;
        while (count < p.AS_tail) {
            if (have && std::fabs(last - p.U) <= 0) count += 1;
            auto sU = [&](int k, double dtk) {
                const double ts = k == 0 ? t : t + dtk;
                last = orc_fermi(ts, p.AS_midpoint, p.AS_rate) * p.U;
                return last;
            };
            auto sQ = [&](int, double, double* w, int* nq) { *w = 0.0; *nq = 0; };
            step(p, W, G1, S, p.AS_dt, sU, sQ);
            have = true;
            t += p.AS_dt;
            ++n;
        }
        if (as_steps) *as_steps = n;
    }
    const double tmax = double(p.time_steps) * p.dt;
    double t = 0.0;
    int64_t n = 0, rec = 0;
    int rc = 0;
    const int64_t n2 = W.n2;
    while (t < tmax) {
        if (dipole_src && n < dens_cap)
            for (int i = 0; i < W.N; ++i) dipole_src[n * W.N + i] = G1[(int64_t)i * (W.N + 1)].imag();
        if (traj && traj_every > 0 && n % traj_every == 0 && rec < traj_cap)
            std::memcpy(traj + 2 * n2 * rec++, G1_, sizeof(cplx) * n2);
        ++n;
        const double dens = 2 * G1[0].imag();
        if (std::fabs(dens) > 3 || std::isnan(dens)) {
            rc = 1;
            break;
        }
        auto sU = [&](int, double) { return p.U; };
        auto sQ = [&](int k, double dtk, double* w, int* nq) {
            const double ts = k == 0 ? t : t + dtk;
            if (ts < p.gate_time) {
                *w = 0.0;
                *nq = 0;
                return;
            }
            const double on = orc_fermi(ts, p.quench_on, p.quench_rate);
            *w = p.q_pulse ? (on * (1 - orc_fermi(ts, p.quench_off, p.quench_rate))) * p.quench_strength
                           : on * p.quench_strength;
            *nq = p.Nq;
        };
        step(p, W, G1, S, p.dt, sU, sQ);
        t += p.dt;
    }
    if (rk_steps) *rk_steps = n;
    W.to_ref(S[0], reinterpret_cast<cplx*>(Gxx_));
    W.to_ref(S[1], reinterpret_cast<cplx*>(Gxy_));
    return rc;
}

}  // extern "C"
