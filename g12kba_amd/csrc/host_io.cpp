// Host-side pieces of the drop-in boundary: parameters and the reference
// `input` format, the non-interacting initial state, observables and text
// outputs, checkpoints.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <algorithm>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/g12kba.h"
#include "host_io.h"

typedef std::complex<double> cplx;

namespace g12io {

// Cyclic Jacobi for a complex Hermitian matrix.  Each rotation first removes
// the phase of a_pq (D = diag(1, conj(e^{i phi}))) and then applies the real
// symmetric Jacobi rotation of the resulting 2x2 block.
int hermitian_eigen(double* a_, int n, double* w, double* v_) {
    cplx* A = reinterpret_cast<cplx*>(a_);
    cplx* V = reinterpret_cast<cplx*>(v_);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[(size_t)i * n + j] = (i == j) ? 1.0 : 0.0;
    double norm = 0.0;
    for (int i = 0; i < n * n; ++i) norm += std::norm(A[i]);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += std::norm(A[(size_t)p * n + q]);
        if (off <= 1e-32 * norm || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const cplx apq = A[(size_t)p * n + q];
                const double mag = std::abs(apq);
                if (mag == 0.0) continue;
                const double app = A[(size_t)p * n + p].real(), aqq = A[(size_t)q * n + q].real();
                const cplx ph = apq / mag;
                const double zeta = (aqq - app) / (2.0 * mag);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                // U = [[c, s], [-s conj(ph), c conj(ph)]] on coordinates (p, q)
                const cplx upp = c, upq = s, uqp = -s * std::conj(ph), uqq = c * std::conj(ph);
                for (int k = 0; k < n; ++k) {  // A <- A U, V <- V U
                    cplx& akp = A[(size_t)k * n + p];
                    cplx& akq = A[(size_t)k * n + q];
                    const cplx x = akp, y = akq;
                    akp = x * upp + y * uqp;
                    akq = x * upq + y * uqq;
                    cplx& vkp = V[(size_t)k * n + p];
                    cplx& vkq = V[(size_t)k * n + q];
                    const cplx vx = vkp, vy = vkq;
                    vkp = vx * upp + vy * uqp;
                    vkq = vx * upq + vy * uqq;
                }
                for (int k = 0; k < n; ++k) {  // A <- U^dagger A
                    cplx& apk = A[(size_t)p * n + k];
                    cplx& aqk = A[(size_t)q * n + k];
                    const cplx x = apk, y = aqk;
                    apk = std::conj(upp) * x + std::conj(uqp) * y;
                    aqk = std::conj(upq) * x + std::conj(uqq) * y;
                }
                A[(size_t)p * n + q] = 0.0;
                A[(size_t)q * n + p] = 0.0;
                A[(size_t)p * n + p] = A[(size_t)p * n + p].real();
                A[(size_t)q * n + q] = A[(size_t)q * n + q].real();
            }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::vector<double> ev(n);
    for (int i = 0; i < n; ++i) ev[i] = A[(size_t)i * n + i].real();
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return ev[x] < ev[y]; });
    std::vector<cplx> Vs((size_t)n * n);
    for (int k = 0; k < n; ++k) {
        w[k] = ev[idx[k]];
        for (int i = 0; i < n; ++i) Vs[(size_t)i * n + k] = V[(size_t)i * n + idx[k]];
    }
    std::memcpy(V, Vs.data(), sizeof(cplx) * n * n);
    return 0;
}

static const char kMagic[8] = {'G', '1', '2', 'K', 'B', 'A', '0', '1'};

int save_checkpoint(const char* path, int Ns, int Nb, double t, int64_t steps, const double* G1, const double* Gxx,
                    const double* Gxy) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return G12_E_IO;
    const int64_t N = (int64_t)Ns * Nb, n2 = N * N, n4 = n2 * n2;
    int32_t dims[2] = {Ns, Nb};
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(dims, sizeof dims, 1, f) == 1 &&
              std::fwrite(&t, sizeof t, 1, f) == 1 && std::fwrite(&steps, sizeof steps, 1, f) == 1 &&
              std::fwrite(G1, sizeof(double) * 2, n2, f) == (size_t)n2 &&
              std::fwrite(Gxx, sizeof(double) * 2, n4, f) == (size_t)n4 &&
              std::fwrite(Gxy, sizeof(double) * 2, n4, f) == (size_t)n4;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : G12_E_IO;
}

int load_checkpoint(const char* path, int Ns, int Nb, double* t, int64_t* steps, double* G1, double* Gxx,
                    double* Gxy) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return G12_E_IO;
    char magic[8];
    int32_t dims[2];
    const int64_t N = (int64_t)Ns * Nb, n2 = N * N, n4 = n2 * n2;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
              std::fread(dims, sizeof dims, 1, f) == 1 && dims[0] == Ns && dims[1] == Nb &&
              std::fread(t, sizeof *t, 1, f) == 1 && std::fread(steps, sizeof *steps, 1, f) == 1 &&
              std::fread(G1, sizeof(double) * 2, n2, f) == (size_t)n2 &&
              std::fread(Gxx, sizeof(double) * 2, n4, f) == (size_t)n4 &&
              std::fread(Gxy, sizeof(double) * 2, n4, f) == (size_t)n4;
    std::fclose(f);
    return ok ? 0 : G12_E_IO;
}

}  // namespace g12io

namespace {

// Value of `key` in a reference-format input file: the line after the first
// line that contains `key` as a substring (read_var_val, read_input.h:53-83).
// The file is scanned twice with std::getline exactly as there, so the same
// corner cases follow (a missing key or value reads as ""; a missing key in
// a file without a final newline reads as its last line).  *found reports
// whether the key occurs at all.
std::string lookup(const std::string& file, const std::string& key, bool* found) {
    std::ifstream f1(file);
    int line_counter = 0;
    std::string line;
    bool hit = false;
    while (std::getline(f1, line)) {
        line_counter++;
        if (line.find(key) != std::string::npos) {
            hit = true;
            break;
        }
    }
    f1.close();
    std::ifstream f2(file);
    std::string s;
    for (int i = 0; i < line_counter + 1; ++i) std::getline(f2, s);
    if (found) *found = hit;
    return s;
}

std::string value_of(const std::string& file, const std::string& key) {
    std::string s = lookup(file, key, nullptr);
    if (s.empty()) {
        std::cout << "Missing value for " << key << " initialising to 0\n";
        return "0";
    }
    return s;
}

}  // namespace

extern "C" {

int g12_params_default(g12_params* p) {
    if (!p) return G12_E_ARG;
    std::memset(p, 0, sizeof *p);
    p->Ns = 2;
    p->Nb = 1;
    p->t1 = 1;
    p->t2 = 0;
    p->t3 = 0;
    p->U = .5;
    p->time_steps = 1000;
    p->dt = .02;
    p->quench_strength = 1;
    p->Nq = 0;
    p->q_type = G12_QUENCH_FULL;
    p->HF = 0;
    p->EHM = 0;
    p->AS_rate = 3;
    p->AS_midpoint = 25;
    p->quench_rate = .2;
    p->quench_on_midpoint = 50;
    p->quench_off_midpoint = 55;
    p->decay_rate = .7;
    p->V_law = G12_VLAW_EXP;
    p->nV_profile = 0;
    p->AS_dt = 0.1;
    p->AS_tail = 20;
    p->quench_gate_time = 50;
    p->p_number = -1;
    p->precision = G12_PREC_FP64;
    return 0;
}

// assign_vals, read_input.h:85-146, plus the optional keys long_range_law,
// filling and precision.
int g12_parse_input(const char* path, g12_params* p) {
    if (!path || !p) return G12_E_ARG;
    const std::string f(path);
    try {
        p->t1 = std::stod(value_of(f, "t1"));
        p->t2 = std::stod(value_of(f, "t2"));
        p->t3 = std::stod(value_of(f, "t3"));
        p->U = std::stod(value_of(f, "U"));
        p->time_steps = std::stoi(value_of(f, "time_steps"));
        p->dt = std::stod(value_of(f, "dt"));
        p->quench_strength = std::stod(value_of(f, "quench_strength"));
        p->Ns = std::stoi(value_of(f, "Ns"));
        p->Nb = std::stoi(value_of(f, "Nb"));
        p->Nq = std::stoi(value_of(f, "Nq"));
        p->AS_rate = std::stod(value_of(f, "AS_rate"));
        p->AS_midpoint = std::stod(value_of(f, "AS_midpoint"));
        p->quench_on_midpoint = std::stod(value_of(f, "quench_on_midpoint"));
        p->quench_off_midpoint = std::stod(value_of(f, "quench_off_midpoint"));
        p->quench_rate = std::stod(value_of(f, "quench_rate"));
    } catch (const std::exception&) {
        return G12_E_ARG;
    }
    const std::string hf = value_of(f, "HF");
    if (hf == "true")
        p->HF = 1;
    else if (hf == "false")
        p->HF = 0;
    else {
        std::cout << "Invalid specification of HF calculation parameter, performing full HF-GKBA calculation.\n";
        p->HF = 0;
    }
    const std::string ehm = value_of(f, "EHM");
    if (ehm == "true")
        p->EHM = 1;
    else if (ehm == "false")
        p->EHM = 0;
    else {
        std::cout << "Invalid specification for EHM parameter. performing full calculation for onsite model.\n";
        p->EHM = 0;
    }
    const std::string qt = value_of(f, "q_type");
    if (qt == "full")
        p->q_type = G12_QUENCH_FULL;
    else if (qt == "pulse")
        p->q_type = G12_QUENCH_PULSE;
    else if (qt == "none") {
        p->quench_strength = 0;
        p->Nq = 0;
    } else {
        std::cout << "Invalid specification of quench type, no quench initiated\n";
    }
    bool found = false;
    std::string v = lookup(f, "long_range_law", &found);
    if (found) {
        if (v == "coulomb")
            p->V_law = G12_VLAW_COULOMB;
        else if (v == "exp")
            p->V_law = G12_VLAW_EXP;
        else
            return G12_E_ARG;
    }
    v = lookup(f, "filling", &found);
    if (found) {
        try {
            p->p_number = std::stoi(v);
        } catch (const std::exception&) {
            return G12_E_ARG;
        }
    }
    v = lookup(f, "precision", &found);
    if (found) {
        if (v == "fp64")
            p->precision = G12_PREC_FP64;
        else if (v == "fp32_3m")
            p->precision = G12_PREC_FP32_3M;
        else if (v == "bf16x3")
            p->precision = G12_PREC_BF16X3;
        else
            return G12_E_ARG;
    }
    return 0;
}

// init_G_test, init_G.h:78-86 (whose stray backslashes do not compile,
// reference defect B-1): every orbital of the first int(Ns/2) sites filled,
// G1[X][X] = i for site(X) < int(Ns/2), all other elements 0.
int g12_init_g_test(const g12_params* p, double* G1) {
    if (!p || !G1 || p->Ns < 1 || p->Nb < 1 || p->Nb > G12_MAX_NB) return G12_E_ARG;
    const int Ns = p->Ns, Nb = p->Nb, N = Ns * Nb;
    cplx* g = reinterpret_cast<cplx*>(G1);
    for (int64_t e = 0; e < (int64_t)N * N; ++e) g[e] = 0.0;
    for (int i = 0; i < Ns / 2; ++i)
        for (int a = 0; a < Nb; ++a) g[(int64_t)(i + a * Ns) * N + i + a * Ns] = cplx(0.0, 1.0);
    return 0;
}

// make_G0, init_G.h:57-75: h(U=0) = eps - t1 (neighbouring sites) - t2
// (adjacent orbitals); G1 = i sum_{k < p} phi_k phi_k^dagger.
int g12_make_g0(const g12_params* p, double* G1, int fp32_emulation) {
    if (!p || !G1 || p->Ns < 1 || p->Nb < 1 || p->Nb > G12_MAX_NB) return G12_E_ARG;
    const int Ns = p->Ns, Nb = p->Nb, N = Ns * Nb;
    const int pn = p->p_number >= 0 ? p->p_number : Ns / 2;
    if (pn > N) return G12_E_ARG;
    std::vector<cplx> h((size_t)N * N, 0.0), v((size_t)N * N);
    std::vector<double> w(N);
    for (int i = 0; i < Ns; ++i)
        for (int a = 0; a < Nb; ++a) {
            const int X = i + a * Ns;
            h[(size_t)X * N + X] += p->eps[a];
            if (i + 1 < Ns) {
                h[(size_t)(X + 1) * N + X] -= p->t1;
                h[(size_t)X * N + X + 1] -= p->t1;
            }
            if (a + 1 < Nb) {
                h[(size_t)X * N + X + Ns] -= p->t2;
                h[(size_t)(X + Ns) * N + X] -= p->t2;
            }
        }
    if (fp32_emulation)
        for (auto& z : h) z = cplx((float)z.real(), (float)z.imag());
    g12io::hermitian_eigen(reinterpret_cast<double*>(h.data()), N, w.data(), reinterpret_cast<double*>(v.data()));
    if (fp32_emulation)
        for (auto& z : v) z = cplx((float)z.real(), (float)z.imag());
    cplx* g = reinterpret_cast<cplx*>(G1);
    const cplx I1(0.0, 1.0);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            cplx acc = 0.0;
            for (int k = 0; k < pn; ++k) acc += v[(size_t)i * N + k] * std::conj(v[(size_t)j * N + k]) * I1;
            g[(size_t)i * N + j] = acc;
        }
    return 0;
}

double g12_fermi(double t, double t0, double tau) { return 1.0 / (1 + std::exp((t0 - t) / tau)); }

// dipole.h:25-33, same summation order (outer loop over the orbital index).
int g12_dipole(double* p, const double* traj, int64_t T, int N) {
    if (!p || (!traj && T > 0) || N < 1) return G12_E_ARG;
    const cplx* G = reinterpret_cast<const cplx*>(traj);
    const int64_t n2 = (int64_t)N * N;
    const double sn = std::sqrt((double)n2);
    for (int64_t t = 0; t < T; ++t) p[t] = 0.0;
    for (int i = 0; i < int(sn / 2); ++i)
        for (int64_t t = 0; t < T; ++t) {
            const cplx* g = G + t * n2;
            p[t] += 2 * (sn / 2 - i - .5) * (g[(int64_t)i * (N + 1)].imag() - g[n2 - (int64_t)i * (N + 1) - 1].imag()) / sn;
        }
    return 0;
}

// write_to_cfile2D / write_to_rfile, write_to_file.h:26-69: default ostream
// formatting, a space after every value, one row per matrix element.
int g12_write_outputs(const char* dir, const double* traj, int64_t T, int N, const double* dipole, double progress) {
    if (!dir) return G12_E_ARG;
    const std::string d(dir);
    const cplx* G = reinterpret_cast<const cplx*>(traj);
    const int64_t n2 = (int64_t)N * N;
    std::ofstream fi(d + "/imag_G1.txt"), fr(d + "/real_G1.txt");
    if (!fi || !fr) return G12_E_IO;
    if (T > 0)
        for (int64_t j = 0; j < n2; ++j) {
            for (int64_t i = 0; i < T; ++i) {
                fi << G[i * n2 + j].imag() << " ";
                fr << G[i * n2 + j].real() << " ";
            }
            fi << "\n";
            fr << "\n";
        }
    if (dipole) {
        std::ofstream fd(d + "/dipole.txt");
        if (!fd) return G12_E_IO;
        for (int64_t i = 0; i < T; ++i) fd << dipole[i] << '\n';
    }
    if (progress >= 0) {
        std::ofstream fp(d + "/progress.txt");
        if (!fp) return G12_E_IO;
        fp << progress << '\n';
    }
    return 0;
}

}  // extern "C"
