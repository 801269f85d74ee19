// C ABI: contexts, device-resident adiabatic switching and RK4 drivers, and
// the routine-level entry points (one per reference routine).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "g12_internal.h"
#include "host_io.h"

using g12::DevState;
using g12::StageParams;

namespace {

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "g12kba: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return G12_E_HIP;                                                          \
        }                                                                              \
    } while (0)

int select_device() {
    static int dev = -1;
    if (dev >= 0) return dev;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return -1;
    const char* s = std::getenv("G12KBA_DEVICE");
    dev = s ? std::atoi(s) : 0;
    if (dev < 0 || dev >= count) dev = 0;
    if (hipSetDevice(dev) != hipSuccess) return -1;
    return dev;
}

// V_n = U_eff * profile_n, n = 1..nV (nV = 0 without long-range interactions).
int v_profile(const g12_params& p, double* prof) {
    if (!p.EHM) return 0;
    int n = p.Ns - 1;
    if (n > G12_MAX_SHELLS) n = G12_MAX_SHELLS;
    if (p.V_law == G12_VLAW_EXPLICIT) {
        if (p.nV_profile < n) n = p.nV_profile;
        for (int i = 0; i < n; ++i) prof[i] = p.V_profile[i];
        return n;
    }
    for (int i = 0; i < n; ++i) {
        const double d = i + 1;
        prof[i] = p.V_law == G12_VLAW_COULOMB ? 1.0 / (1.0 + d) : std::exp(-p.decay_rate * d);
    }
    return n;
}

bool correlated(const g12_params& p) { return p.U != 0.0 && !p.HF; }

StageParams stage_params(const g12_params& p, const double* prof, int nV, int k, double dt_full, double U_eff,
                         double wq, int Nq) {
    StageParams sp{};
    const bool half = (k == 1 || k == 2);
    sp.c = half ? dt_full / 2 : dt_full;
    sp.wk = half ? 2.0 : 1.0;
    sp.dt = dt_full;
    sp.first = k == 0;
    sp.last = k == 3;
    sp.corr = correlated(p);
    sp.U = U_eff;
    sp.nV = nV;
    for (int i = 0; i < nV; ++i) sp.V[i] = U_eff * prof[i];
    sp.wq = wq;
    sp.Nq = Nq;
    sp.t1 = p.t1;
    sp.t2 = p.t2;
    for (int i = 0; i < G12_MAX_NB; ++i) sp.eps[i] = i < p.Nb ? p.eps[i] : 0.0;
    return sp;
}

struct Timing {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[3];
    int64_t count[3] = {0, 0, 0};
    double ms[3] = {0, 0, 0};
    std::vector<hipEvent_t> pool;

    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    void resolve() {
        for (int c = 0; c < 3; ++c) {
            for (auto& pr : pending[c]) {
                float t = 0.f;
                hipEventSynchronize(pr.second);
                if (hipEventElapsedTime(&t, pr.first, pr.second) == hipSuccess) {
                    ms[c] += t;
                    count[c] += 1;
                }
                pool.push_back(pr.first);
                pool.push_back(pr.second);
            }
            pending[c].clear();
        }
    }
    void destroy() {
        resolve();
        for (auto e : pool) hipEventDestroy(e);
        pool.clear();
    }
};

}  // namespace

struct g12_ctx {
    g12_params p;
    DevState d;
    hipStream_t stream = nullptr;
    double prof[G12_MAX_SHELLS];
    int nV = 0;
    double t = 0.0;      // production time
    int64_t steps = 0;   // production iterations begun
    Timing timing;
    int64_t bytes = 0;
    std::vector<void*> allocs;

    template <class T>
    int alloc(T** ptr, int64_t count) {
        void* v = nullptr;
        if (count <= 0) count = 1;
        if (hipMalloc(&v, sizeof(T) * count) != hipSuccess) return G12_E_NOMEM;
        allocs.push_back(v);
        bytes += sizeof(T) * count;
        *ptr = static_cast<T*>(v);
        return 0;
    }
};

namespace {

template <class F>
void timed(g12_ctx* c, int cls, F&& launch) {
    if (!c->timing.on) {
        launch();
        return;
    }
    hipEvent_t a = c->timing.get(), b = c->timing.get();
    hipEventRecord(a, c->stream);
    launch();
    hipEventRecord(b, c->stream);
    c->timing.pending[cls].push_back({a, b});
    if (c->timing.pending[cls].size() > 4096) c->timing.resolve();
}

// One RK4 stage on the device.
void run_stage(g12_ctx* c, const StageParams& sp) {
    const int prec = c->p.precision;
    hipStream_t s = c->stream;
    timed(c, 2, [&] { g12::launch_stage_g1(c->d, sp, s); });
    if (sp.corr) {
        timed(c, 0, [&] { g12::launch_pass_strided(c->d, sp, prec, s); });
        timed(c, 2, [&] { g12::launch_tables(c->d, sp, s); });
    }
    timed(c, 2, [&] { g12::launch_g1_rhs(c->d, sp, s); });
    if (sp.corr) timed(c, 1, [&] { g12::launch_pass_fused(c->d, sp, prec, s); });
}

// Quench shift and extent for a production stage at time ts (RK_solver.h:141-151).
void quench_at(const g12_params& p, double ts, double* w, int* nq) {
    if (ts < p.quench_gate_time) {
        *w = 0.0;
        *nq = 0;
        return;
    }
    const double on = g12_fermi(ts, p.quench_on_midpoint, p.quench_rate);
    if (p.q_type == G12_QUENCH_PULSE)
        *w = (on * (1 - g12_fermi(ts, p.quench_off_midpoint, p.quench_rate))) * p.quench_strength;
    else
        *w = on * p.quench_strength;
    *nq = p.Nq;
}

void production_step(g12_ctx* c) {
    const g12_params& p = c->p;
    for (int k = 0; k < 4; ++k) {
        const double dtk = (k == 1 || k == 2) ? p.dt / 2 : p.dt;
        const double ts = k == 0 ? c->t : c->t + dtk;
        double wq;
        int nq;
        quench_at(p, ts, &wq, &nq);
        run_stage(c, stage_params(p, c->prof, c->nV, k, p.dt, p.U, wq, nq));
    }
    c->t += p.dt;
}

int read_fail(g12_ctx* c, int* fail) {
    HIPCHK(hipMemcpyAsync(fail, c->d.fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

bool valid_geometry(int Ns, int Nb) { return Ns >= 1 && Nb >= 1 && Nb <= G12_MAX_NB && Ns * Nb <= 64; }

}  // namespace

extern "C" {

const char* g12_version(void) { return "g12kba-mi355x 0.1"; }

int g12_create(const g12_params* p, g12_ctx** out) {
    if (!p || !out) return G12_E_ARG;
    if (!valid_geometry(p->Ns, p->Nb)) return G12_E_SIZE;
    if (select_device() < 0) return G12_E_HIP;
    g12_ctx* c = new (std::nothrow) g12_ctx();
    if (!c) return G12_E_NOMEM;
    c->p = *p;
    c->nV = v_profile(c->p, c->prof);
    DevState& d = c->d;
    d.Ns = p->Ns;
    d.Nb = p->Nb;
    d.N = p->Ns * p->Nb;
    d.n2 = (int64_t)d.N * d.N;
    d.n3 = d.n2 * d.N;
    d.n4 = d.n2 * d.n2;
    int rc = 0;
    rc |= c->alloc(&d.S1, d.n2);
    rc |= c->alloc(&d.K1, d.n2);
    rc |= c->alloc(&d.A1, d.n2);
    rc |= c->alloc(&d.Y1, d.n2);
    rc |= c->alloc(&d.h, d.n2);
    rc |= c->alloc(&d.I, d.n2);
    rc |= c->alloc(&d.Wxy, d.n2);
    rc |= c->alloc(&d.Wxx, d.n2);
    const bool corr = correlated(*p);
    const int64_t n4 = corr ? d.n4 : 1, n3 = corr ? d.n3 : 1;
    for (int s = 0; s < 2; ++s) {
        rc |= c->alloc(&d.S2[s], d.n4);  // state is kept even when frozen
        rc |= c->alloc(&d.K2[s], n4);
        rc |= c->alloc(&d.A2[s], n4);
        rc |= c->alloc(&d.R2[s], n4);
        rc |= c->alloc(&d.D[s], n3);
        rc |= c->alloc(&d.F[s], n3);
        rc |= c->alloc(&d.Ft[s], n3);
        rc |= c->alloc(&d.Q[s], n3);
        rc |= c->alloc(&d.Qp[s], n3);
    }
    rc |= c->alloc(&d.fail, 1);
    d.traj_slots = 256;
    rc |= c->alloc(&d.traj, d.n2 * d.traj_slots);
    if (rc) {
        g12_destroy(c);
        return G12_E_NOMEM;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        g12_destroy(c);
        return G12_E_HIP;
    }
    hipMemsetAsync(d.fail, 0, sizeof(int), c->stream);
    hipMemsetAsync(d.S1, 0, sizeof(double2) * d.n2, c->stream);
    hipMemsetAsync(d.I, 0, sizeof(double2) * d.n2, c->stream);
    for (int s = 0; s < 2; ++s) hipMemsetAsync(d.S2[s], 0, sizeof(double2) * d.n4, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        g12_destroy(c);
        return G12_E_HIP;
    }
    *out = c;
    return 0;
}

int g12_destroy(g12_ctx* c) {
    if (!c) return 0;
    if (c->stream) hipStreamSynchronize(c->stream);
    c->timing.destroy();
    for (void* v : c->allocs) hipFree(v);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int g12_synchronize(void) {
    if (select_device() < 0) return G12_E_HIP;
    return hipDeviceSynchronize() == hipSuccess ? 0 : G12_E_HIP;
}

int g12_memory_bytes(g12_ctx* c, int64_t* bytes) {
    if (!c || !bytes) return G12_E_ARG;
    *bytes = c->bytes;
    return 0;
}

int g12_set_state(g12_ctx* c, const double* G1, const double* G2xx, const double* G2xy) {
    if (!c || !G1) return G12_E_ARG;
    DevState& d = c->d;
    HIPCHK(hipMemcpyAsync(d.S1, G1, sizeof(double2) * d.n2, hipMemcpyHostToDevice, c->stream));
    const double* g2[2] = {G2xx, G2xy};
    double2* tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, sizeof(double2) * d.n4));
    for (int s = 0; s < 2; ++s) {
        if (!g2[s]) {
            hipMemsetAsync(d.S2[s], 0, sizeof(double2) * d.n4, c->stream);
            continue;
        }
        hipMemcpyAsync(tmp, g2[s], sizeof(double2) * d.n4, hipMemcpyHostToDevice, c->stream);
        g12::launch_ref_to_nat(tmp, d.S2[s], d.Ns, d.Nb, c->stream);
    }
    hipMemsetAsync(d.fail, 0, sizeof(int), c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    hipFree(tmp);
    if (e != hipSuccess) return G12_E_HIP;
    c->t = 0.0;
    c->steps = 0;
    return 0;
}

int g12_get_state(g12_ctx* c, double* G1, double* G2xx, double* G2xy) {
    if (!c) return G12_E_ARG;
    DevState& d = c->d;
    if (G1) HIPCHK(hipMemcpyAsync(G1, d.S1, sizeof(double2) * d.n2, hipMemcpyDeviceToHost, c->stream));
    double* g2[2] = {G2xx, G2xy};
    double2* tmp = nullptr;
    if (G2xx || G2xy) HIPCHK(hipMalloc(&tmp, sizeof(double2) * d.n4));
    for (int s = 0; s < 2; ++s) {
        if (!g2[s]) continue;
        g12::launch_nat_to_ref(d.S2[s], tmp, d.Ns, d.Nb, c->stream);
        hipMemcpyAsync(g2[s], tmp, sizeof(double2) * d.n4, hipMemcpyDeviceToHost, c->stream);
    }
    hipError_t e = hipStreamSynchronize(c->stream);
    if (tmp) hipFree(tmp);
    return e == hipSuccess ? 0 : G12_E_HIP;
}

int g12_get_time(g12_ctx* c, double* t, int64_t* steps) {
    if (!c) return G12_E_ARG;
    if (t) *t = c->t;
    if (steps) *steps = c->steps;
    return 0;
}

// Adiabatic switching (AS.h:63-214): fixed step AS_dt, U_AS = fermi(t) U per
// stage with V scaled alike, no quench, until AS_tail iterations have begun
// with U_AS == U (the first iteration never counts).
int g12_adiabatic_switching(g12_ctx* c, int* steps) {
    if (!c) return G12_E_ARG;
    const g12_params& p = c->p;
    int count = 0, n = 0;
    bool have = false;
    double last = 0.0, t = 0.0;
    while (count < p.AS_tail) {
        if (have && std::fabs(last - p.U) <= 0) count += 1;
        for (int k = 0; k < 4; ++k) {
            const double dtk = (k == 1 || k == 2) ? p.AS_dt / 2 : p.AS_dt;
            const double ts = k == 0 ? t : t + dtk;
            last = g12_fermi(ts, p.AS_midpoint, p.AS_rate) * p.U;
            run_stage(c, stage_params(p, c->prof, c->nV, k, p.AS_dt, last, 0.0, 0));
        }
        have = true;
        t += p.AS_dt;
        ++n;
        if (n > 100000000) break;  // unreachable for finite parameters
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (steps) *steps = n;
    return 0;
}

// Production RK4 (RK_solver.h:64-226).
int g12_rk4(g12_ctx* c, double* traj, int64_t traj_cap, int64_t* steps) {
    if (!c) return G12_E_ARG;
    const g12_params& p = c->p;
    const double tmax = double(p.time_steps) * p.dt;
    DevState& d = c->d;
    c->t = 0.0;
    c->steps = 0;
    HIPCHK(hipMemsetAsync(d.fail, 0, sizeof(int), c->stream));
    int64_t n = 0, flushed = 0;
    int fail = 0;
    auto flush = [&]() -> int {
        // copy recorded slots [flushed, n) to the host trajectory
        if (traj && flushed < n && flushed < traj_cap) {
            const int64_t hi = n < traj_cap ? n : traj_cap;
            for (int64_t i = flushed; i < hi; ++i) {
                const int slot = (int)(i % d.traj_slots);
                HIPCHK(hipMemcpyAsync(traj + i * d.n2 * 2, d.traj + slot * d.n2, sizeof(double2) * d.n2,
                                      hipMemcpyDeviceToHost, c->stream));
            }
        }
        flushed = n;
        return read_fail(c, &fail);
    };
    while (c->t < tmax) {
        g12::launch_record(d, (int)(n % d.traj_slots), n, c->stream);
        ++n;
        production_step(c);
        if (n - flushed >= d.traj_slots) {
            int rc = flush();
            if (rc) return rc;
            if (fail) break;
        }
    }
    int rc = flush();
    if (rc) return rc;
    c->timing.resolve();
    if (fail) {
        c->steps = fail;
        if (steps) *steps = fail;
        return G12_GUARD;
    }
    c->steps = n;
    if (steps) *steps = n;
    return 0;
}

int g12_rk4_advance(g12_ctx* c, int64_t nsteps, int64_t* done) {
    if (!c || nsteps < 0) return G12_E_ARG;
    DevState& d = c->d;
    int fail = 0;
    const double2* keep = d.traj;
    for (int64_t i = 0; i < nsteps; ++i) {
        DevState dd = d;
        dd.traj = nullptr;  // guard only
        g12::launch_record(dd, 0, c->steps, c->stream);
        c->steps += 1;
        production_step(c);
    }
    (void)keep;
    int rc = read_fail(c, &fail);
    if (rc) return rc;
    c->timing.resolve();
    if (done) *done = nsteps;
    return fail ? G12_GUARD : 0;
}

int g12_timing_enable(g12_ctx* c, int on) {
    if (!c) return G12_E_ARG;
    c->timing.resolve();
    c->timing.on = on != 0;
    for (int i = 0; i < 3; ++i) {
        c->timing.count[i] = 0;
        c->timing.ms[i] = 0;
    }
    return 0;
}

int g12_timing_get(g12_ctx* c, int which, int64_t* launches, double* total_ms) {
    if (!c || which < 0 || which > 2) return G12_E_ARG;
    c->timing.resolve();
    if (launches) *launches = c->timing.count[which];
    if (total_ms) *total_ms = c->timing.ms[which];
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// routine-level entry points

namespace {

struct Scratch {
    std::vector<void*> v;
    ~Scratch() {
        for (void* p : v) hipFree(p);
    }
    template <class T>
    T* get(int64_t n) {
        void* p = nullptr;
        if (hipMalloc(&p, sizeof(T) * (n > 0 ? n : 1)) != hipSuccess) return nullptr;
        v.push_back(p);
        return static_cast<T*>(p);
    }
};

hipStream_t api_stream() { return nullptr; }  // legacy default stream: calls are synchronous

StageParams interaction_params(int Ns, int Nb, double U, const double* V, int nV) {
    StageParams sp{};
    sp.U = U;
    sp.nV = nV < G12_MAX_SHELLS ? nV : G12_MAX_SHELLS;
    for (int i = 0; i < sp.nV; ++i) sp.V[i] = V[i];
    (void)Ns;
    (void)Nb;
    return sp;
}

int up(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : G12_E_HIP;
}
int down(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : G12_E_HIP;
}

#define DEV_OR_FAIL(ptr) \
    if (!(ptr)) return G12_E_NOMEM

int check_routine_geometry(int Ns, int Nb, int nV) {
    if (!valid_geometry(Ns, Nb) || nV < 0 || nV > G12_MAX_SHELLS) return G12_E_SIZE;
    if (select_device() < 0) return G12_E_HIP;
    return 0;
}

}  // namespace

extern "C" {

int g12_hf_hamiltonian(double* h, const double* G1x, const double* G1y, int Ns, int Nb, double U, double t1,
                       double t2, double w, const double* eps, const double* V, int nV, int quench_extent) {
    (void)G1y;
    if (!h || !G1x || (nV > 0 && !V)) return G12_E_ARG;
    int rc = check_routine_geometry(Ns, Nb, nV);
    if (rc) return rc;
    const int64_t n2 = (int64_t)Ns * Nb * Ns * Nb;
    Scratch sc;
    double2* dh = sc.get<double2>(n2);
    double2* dg = sc.get<double2>(n2);
    DEV_OR_FAIL(dh && dg);
    StageParams sp = interaction_params(Ns, Nb, U, V, nV);
    sp.t1 = t1;
    sp.t2 = t2;
    sp.wq = w;
    sp.Nq = quench_extent;
    for (int i = 0; i < Nb; ++i) sp.eps[i] = eps ? eps[i] : 0.0;
    if ((rc = up(dh, h, sizeof(double2) * n2)) || (rc = up(dg, G1x, sizeof(double2) * n2))) return rc;
    g12::launch_hf(dh, dg, Ns, Nb, sp, api_stream());
    return down(h, dh, sizeof(double2) * n2);
}

int g12_h1_commutator(const double* G, const double* h, double* comm, int N) {
    if (!G || !h || !comm || N < 1 || N > 4096) return G12_E_ARG;
    if (select_device() < 0) return G12_E_HIP;
    const int64_t n2 = (int64_t)N * N;
    Scratch sc;
    double2 *dg = sc.get<double2>(n2), *dh = sc.get<double2>(n2), *dc = sc.get<double2>(n2);
    DEV_OR_FAIL(dg && dh && dc);
    int rc;
    if ((rc = up(dg, G, sizeof(double2) * n2)) || (rc = up(dh, h, sizeof(double2) * n2))) return rc;
    g12::launch_comm1(dg, dh, dc, N, api_stream());
    return down(comm, dc, sizeof(double2) * n2);
}

int g12_h2_commutator(const double* G2, const double* h1x, const double* h1y, double* comm, int Ns, int Nb,
                      int precision) {
    if (!G2 || !h1x || !h1y || !comm) return G12_E_ARG;
    int rc = check_routine_geometry(Ns, Nb, 0);
    if (rc) return rc;
    const int N = Ns * Nb;
    const int64_t n2 = (int64_t)N * N, n4 = n2 * n2;
    Scratch sc;
    double2 *draw = sc.get<double2>(n4), *dT = sc.get<double2>(n4), *dC = sc.get<double2>(n4);
    double2 *dhx = sc.get<double2>(n2), *dhy = sc.get<double2>(n2);
    DEV_OR_FAIL(draw && dT && dC && dhx && dhy);
    if ((rc = up(draw, G2, sizeof(double2) * n4)) || (rc = up(dhx, h1x, sizeof(double2) * n2)) ||
        (rc = up(dhy, h1y, sizeof(double2) * n2)))
        return rc;
    g12::launch_ref_to_nat(draw, dT, Ns, Nb, api_stream());
    g12::launch_comm2(dT, dhx, dhy, dC, nullptr, N, precision, api_stream());
    g12::launch_nat_to_ref(dC, draw, Ns, Nb, api_stream());
    return down(comm, draw, sizeof(double2) * n4);
}

int g12_collision(double* Ix, const double* G2xyxy, const double* G2xxxx, int Ns, int Nb, double U,
                  const double* V, int nV) {
    if (!Ix || !G2xyxy || !G2xxxx || (nV > 0 && !V)) return G12_E_ARG;
    int rc = check_routine_geometry(Ns, Nb, nV);
    if (rc) return rc;
    const int N = Ns * Nb;
    const int64_t n2 = (int64_t)N * N, n3 = n2 * N, n4 = n2 * n2;
    Scratch sc;
    double2 *draw = sc.get<double2>(n4), *dT = sc.get<double2>(n4);
    double2 *dP = sc.get<double2>(n3), *dQ = sc.get<double2>(n3), *dF = sc.get<double2>(n3);
    double2* dI = sc.get<double2>(n2);
    double *wxy = sc.get<double>(n2), *wxx = sc.get<double>(n2);
    DEV_OR_FAIL(draw && dT && dP && dQ && dF && dI && wxy && wxx);
    StageParams sp = interaction_params(Ns, Nb, U, V, nV);
    const double* src[2] = {G2xyxy, G2xxxx};
    double2* dst[2] = {dP, dQ};
    for (int s = 0; s < 2; ++s) {
        if ((rc = up(draw, src[s], sizeof(double2) * n4))) return rc;
        g12::launch_ref_to_nat(draw, dT, Ns, Nb, api_stream());
        g12::launch_d_gather(dT, dst[s], N, api_stream());
    }
    g12::launch_w_tables(wxy, wxx, Ns, Nb, sp, api_stream());
    g12::launch_f_tables(dP, dQ, wxy, wxx, dF, nullptr, N, api_stream());
    g12::launch_collision_from_f(dF, dI, N, api_stream());
    return down(Ix, dI, sizeof(double2) * n2);
}

}  // extern "C"

static int psi_common(double* psi, const double* G1, int Ns, int Nb, double U, const double* V, int nV,
                      bool same_spin) {
    if (!psi || !G1 || (nV > 0 && !V)) return G12_E_ARG;
    int rc = check_routine_geometry(Ns, Nb, nV);
    if (rc) return rc;
    const int N = Ns * Nb;
    const int64_t n2 = (int64_t)N * N, n3 = n2 * N, n4 = n2 * n2;
    Scratch sc;
    double2 *dg = sc.get<double2>(n2), *dQ = sc.get<double2>(n3), *dQp = sc.get<double2>(n3);
    double2 *dnat = sc.get<double2>(n4), *dref = sc.get<double2>(n4);
    double *wxy = sc.get<double>(n2), *wxx = sc.get<double>(n2);
    DEV_OR_FAIL(dg && dQ && dQp && dnat && dref && wxy && wxx);
    StageParams sp = interaction_params(Ns, Nb, U, V, nV);
    if ((rc = up(dg, G1, sizeof(double2) * n2))) return rc;
    g12::launch_w_tables(wxy, wxx, Ns, Nb, sp, api_stream());
    const double* W = same_spin ? wxx : wxy;
    g12::launch_q_table(dg, W, dQ, dQp, N, api_stream());
    g12::launch_psi(dg, W, dQ, dQp, dnat, N, api_stream());
    g12::launch_nat_to_ref(dnat, dref, Ns, Nb, api_stream());
    return down(psi, dref, sizeof(double2) * n4);
}

extern "C" {

int g12_psi_xxxx(double* psi, const double* G1, int Ns, int Nb, double U, const double* V, int nV) {
    return psi_common(psi, G1, Ns, Nb, U, V, nV, true);
}

int g12_psi_xyxy(double* psi, const double* G1, int Ns, int Nb, double U, const double* V, int nV) {
    return psi_common(psi, G1, Ns, Nb, U, V, nV, false);
}

int g12_pi(double* pi, const double* G1y, const double* G2xyxy, const double* G2xxxx, double U, int Ns, int Nb,
           const double* V, int nV) {
    if (!pi || !G1y || !G2xyxy || !G2xxxx || (nV > 0 && !V)) return G12_E_ARG;
    int rc = check_routine_geometry(Ns, Nb, nV);
    if (rc) return rc;
    const int N = Ns * Nb;
    const int64_t n2 = (int64_t)N * N, n3 = n2 * N, n4 = n2 * n2;
    Scratch sc;
    double2 *draw = sc.get<double2>(n4), *dT = sc.get<double2>(n4);
    double2 *dP = sc.get<double2>(n3), *dQ = sc.get<double2>(n3), *dF = sc.get<double2>(n3);
    double2* dg = sc.get<double2>(n2);
    double *wxy = sc.get<double>(n2), *wxx = sc.get<double>(n2);
    DEV_OR_FAIL(draw && dT && dP && dQ && dF && dg && wxy && wxx);
    StageParams sp = interaction_params(Ns, Nb, U, V, nV);
    const double* src[2] = {G2xyxy, G2xxxx};
    double2* dst[2] = {dP, dQ};
    for (int s = 0; s < 2; ++s) {
        if ((rc = up(draw, src[s], sizeof(double2) * n4))) return rc;
        g12::launch_ref_to_nat(draw, dT, Ns, Nb, api_stream());
        g12::launch_d_gather(dT, dst[s], N, api_stream());
    }
    if ((rc = up(dg, G1y, sizeof(double2) * n2))) return rc;
    g12::launch_w_tables(wxy, wxx, Ns, Nb, sp, api_stream());
    g12::launch_f_tables(dP, dQ, wxy, wxx, dF, nullptr, N, api_stream());
    g12::launch_pi(dg, dF, dT, N, api_stream());
    g12::launch_nat_to_ref(dT, draw, Ns, Nb, api_stream());
    return down(pi, draw, sizeof(double2) * n4);
}

int g12_g1_step(const double* comm, const double* I, double* K, double* RK_result, double w, int Ns, int Nb) {
    if (!comm || !I || !K || !RK_result) return G12_E_ARG;
    int rc = check_routine_geometry(Ns, Nb, 0);
    if (rc) return rc;
    const int N = Ns * Nb;
    const int64_t n2 = (int64_t)N * N;
    Scratch sc;
    double2 *dc = sc.get<double2>(n2), *dI = sc.get<double2>(n2), *dK = sc.get<double2>(n2),
            *dR = sc.get<double2>(n2);
    DEV_OR_FAIL(dc && dI && dK && dR);
    if ((rc = up(dc, comm, sizeof(double2) * n2)) || (rc = up(dI, I, sizeof(double2) * n2)) ||
        (rc = up(dR, RK_result, sizeof(double2) * n2)))
        return rc;
    g12::launch_g1_step(dc, dI, dK, dR, w, N, api_stream());
    if ((rc = down(K, dK, sizeof(double2) * n2))) return rc;
    return down(RK_result, dR, sizeof(double2) * n2);
}

int g12_g2_step(const double* comm, const double* psi, const double* pi, double* K, double* RK_result, double w,
                int Ns, int Nb) {
    if (!comm || !psi || !pi || !K || !RK_result) return G12_E_ARG;
    int rc = check_routine_geometry(Ns, Nb, 0);
    if (rc) return rc;
    const int N = Ns * Nb;
    const int64_t n2 = (int64_t)N * N, n4 = n2 * n2;
    Scratch sc;
    double2* raw = sc.get<double2>(n4);
    double2 *dc = sc.get<double2>(n4), *ds = sc.get<double2>(n4), *dp = sc.get<double2>(n4),
            *dK = sc.get<double2>(n4), *dR = sc.get<double2>(n4);
    DEV_OR_FAIL(raw && dc && ds && dp && dK && dR);
    const double* src[4] = {comm, psi, pi, RK_result};
    double2* dst[4] = {dc, ds, dp, dR};
    for (int i = 0; i < 4; ++i) {
        if ((rc = up(raw, src[i], sizeof(double2) * n4))) return rc;
        g12::launch_ref_to_nat(raw, dst[i], Ns, Nb, api_stream());
    }
    g12::launch_g2_step(dc, ds, dp, dK, dR, w, N, api_stream());
    g12::launch_nat_to_ref(dK, raw, Ns, Nb, api_stream());
    if ((rc = down(K, raw, sizeof(double2) * n4))) return rc;
    g12::launch_nat_to_ref(dR, raw, Ns, Nb, api_stream());
    return down(RK_result, raw, sizeof(double2) * n4);
}

// ---------------------------------------------------------------------------
// checkpoints: magic, N, Ns, Nb, t, steps, G1, G2xx, G2xy (reference layouts)

int g12_save_state(g12_ctx* c, const char* path) {
    if (!c || !path) return G12_E_ARG;
    const DevState& d = c->d;
    std::vector<double> g1(2 * d.n2), gxx(2 * d.n4), gxy(2 * d.n4);
    int rc = g12_get_state(c, g1.data(), gxx.data(), gxy.data());
    if (rc) return rc;
    return g12io::save_checkpoint(path, d.Ns, d.Nb, c->t, c->steps, g1.data(), gxx.data(), gxy.data());
}

int g12_load_state(g12_ctx* c, const char* path) {
    if (!c || !path) return G12_E_ARG;
    const DevState& d = c->d;
    std::vector<double> g1(2 * d.n2), gxx(2 * d.n4), gxy(2 * d.n4);
    double t = 0;
    int64_t steps = 0;
    int rc = g12io::load_checkpoint(path, d.Ns, d.Nb, &t, &steps, g1.data(), gxx.data(), gxy.data());
    if (rc) return rc;
    rc = g12_set_state(c, g1.data(), gxx.data(), gxy.data());
    if (rc) return rc;
    c->t = t;
    c->steps = steps;
    return 0;
}

}  // extern "C"
