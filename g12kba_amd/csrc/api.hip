// C ABI: runs (sets of trajectories advanced together; a context is a run of
// one trajectory), the device-resident adiabatic-switching and RK4 drivers
// with hipGraph replay of whole RK4 steps, the tile-shape autotuner, and the
// routine-level entry points (one per reference routine).
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

#include "g12_internal.h"
#include "host_io.h"

using g12::BatchDev;
using g12::Bucket;
using g12::DevState;
using g12::StageParams;
using g12::StageRef;
using g12::StageShared;
using g12::TrajParams;

namespace {

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "g12kba: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return G12_E_HIP;                                                          \
        }                                                                              \
    } while (0)

// ---- scratch of the no-allocation entry point -------------------------------
// g12_h2_commutator_noalloc keeps its launch descriptors and the natural-layout
// copies of its input and result in this array; every GPU the library runs on
// holds its own copy (at its own address) until the process ends.  Host code
// never uses the array's name as an address: scratch_address() looks the
// device address up once per device, with that device current.
constexpr int64_t NOALLOC_N4 = (int64_t)G12_NOALLOC_MAX_N * G12_NOALLOC_MAX_N * G12_NOALLOC_MAX_N * G12_NOALLOC_MAX_N;
constexpr int64_t NOALLOC_BYTES = g12::COMM2_DESC_BYTES + 2 * NOALLOC_N4 * (int64_t)sizeof(double2);
static_assert(NOALLOC_BYTES == G12_NOALLOC_SCRATCH_BYTES, "scratch size as declared in g12kba.h");
static __device__ __align__(256) char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
void* scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on device dev (which must be current), or
// nullptr with a message when the lookup fails.
void* scratch_address(int dev) {
    if (dev < 0 || dev >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            std::fprintf(stderr, "g12kba: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                         hipGetErrorString(e));
            return nullptr;
        }
        scratch_ptr[dev] = p;
    }
    return scratch_ptr[dev];
}

int select_device() {
    static int dev = -1;
    if (dev >= 0) return dev;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return -1;
    const char* s = std::getenv("G12KBA_DEVICE");
    dev = s ? std::atoi(s) : 0;
    if (dev < 0 || dev >= count) dev = 0;
    if (hipSetDevice(dev) != hipSuccess) return -1;
    // library set-up on the selected device: the no-allocation scratch address
    (void)scratch_address(dev);
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

bool valid_geometry(int Ns, int Nb) { return Ns >= 1 && Nb >= 1 && Nb <= G12_MAX_NB && Ns * Nb <= G12_MAX_N; }

// The mixed-precision tile products hold a whole tile in LDS: N <= 64.
bool valid_precision_size(int N, int precision) { return precision == G12_PREC_FP64 || N <= G12_MAX_N_MIXED; }

// The time protocol (steps, switching, quench ramps) and the precision are
// shared by every trajectory of a run; the physics (sizes, hoppings,
// interactions, quench strength and extent) is per trajectory.
bool same_protocol(const g12_params& a, const g12_params& b) {
    return a.dt == b.dt && a.time_steps == b.time_steps && a.AS_dt == b.AS_dt && a.AS_rate == b.AS_rate &&
           a.AS_midpoint == b.AS_midpoint && a.AS_tail == b.AS_tail && a.quench_rate == b.quench_rate &&
           a.quench_on_midpoint == b.quench_on_midpoint && a.quench_off_midpoint == b.quench_off_midpoint &&
           a.q_type == b.q_type && a.quench_gate_time == b.quench_gate_time && a.precision == b.precision;
}

// Length of the adiabatic switching of a trajectory (AS.h:87-214): the loop
// runs while fewer than AS_tail iterations have begun with U_AS == U exactly,
// U_AS being the value set by the last stage of the previous iteration (the
// first iteration never counts).  Depends on U through the rounding of
// fermi(t) * U.
// Returns -1 when the switching would not end within MAX_AS_STEPS steps (the
// exact test can fail forever, e.g. for a ramp that never reaches 1 in double).
constexpr int MAX_AS_STEPS = 10000000;
int as_step_count(const g12_params& p) {
    int count = 0, n = 0;
    bool have = false;
    double last = 0.0, t = 0.0;
    while (count < p.AS_tail) {
        if (have && std::fabs(last - p.U) <= 0) count += 1;
        for (int k = 0; k < 4; ++k) {
            const double dtk = (k == 1 || k == 2) ? p.AS_dt / 2 : p.AS_dt;
            const double ts = k == 0 ? t : t + dtk;
            last = g12_fermi(ts, p.AS_midpoint, p.AS_rate) * p.U;
        }
        have = true;
        t += p.AS_dt;
        if (++n > MAX_AS_STEPS) return -1;
    }
    return n;
}

// Parameters that would make a driver loop forever or compute with
// non-finite numbers.  The reference's parser gives 0 for a missing key
// (read_input.h:53-83), and those runs end: AS_rate = 0 makes the ramp a step
// (exp of an infinite argument), dt = 0 gives tmax = 0 and no production step.
// A negative AS_rate (a ramp that falls: the exact U_AS == U test of
// AS.h:96-100 never succeeds), AS_dt <= 0 (the switching clock stands still)
// and a production clock that never reaches tmax are rejected.
bool valid_protocol(const g12_params& p) {
    const double v[] = {p.t1, p.t2, p.U, p.dt, p.quench_strength, p.AS_rate, p.AS_midpoint, p.quench_rate,
                        p.quench_on_midpoint, p.quench_off_midpoint, p.decay_rate, p.AS_dt, p.quench_gate_time};
    for (double x : v)
        if (!std::isfinite(x)) return false;
    for (int i = 0; i < G12_MAX_NB; ++i)
        if (!std::isfinite(p.eps[i])) return false;
    if (p.V_law == G12_VLAW_EXPLICIT)
        for (int i = 0; i < p.nV_profile && i < G12_MAX_SHELLS; ++i)
            if (!std::isfinite(p.V_profile[i])) return false;
    const bool clock_ends = p.dt > 0 || double(p.time_steps) * p.dt <= 0;  // while (t < tmax), RK_solver.h:96
    return clock_ends && p.AS_dt > 0 && p.AS_rate >= 0 && p.AS_tail >= 0 && p.Nq >= 0;
}

// Scalars of RK4 stage k of a step of length dt_full.
StageShared stage_block(int k, double dt_full) {
    StageShared ss{};
    const bool half = (k == 1 || k == 2);
    ss.c = half ? dt_full / 2 : dt_full;
    ss.wk = half ? 2.0 : 1.0;
    ss.dt = dt_full;
    ss.cn = (k == 0 || k == 1) ? dt_full / 2 : dt_full;  // input of stage k + 1 (unused at k = 3)
    ss.first = k == 0;
    ss.last = k == 3;
    ss.fold = k == 2;
    ss.lam = 1.0;
    return ss;
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

// Device memory of a run is one allocation carved into 256-byte aligned
// pieces; the first pass (base == nullptr) only measures.
struct Carver {
    char* base;
    int64_t off = 0;
    template <class T>
    T* take(int64_t count) {
        if (count < 1) count = 1;
        off = (off + 255) & ~int64_t(255);
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += (int64_t)sizeof(T) * count;
        return p;
    }
};

}  // namespace

namespace g12 {

// A set of trajectories advanced together.
struct Run {
    int ntraj = 0;
    std::vector<g12_params> p;
    std::vector<DevState> ds;      // device pointers of each trajectory
    std::vector<TrajParams> tp;
    std::vector<int> as_steps;
    int max_as = 0;
    int maxN = 0;
    bool any_corr = false;
    int precision = 0;
    DevState* d_ds = nullptr;
    TrajParams* d_tp = nullptr;
    int* d_fail = nullptr;          // [ntraj]
    int* d_counter = nullptr;       // step within the current window
    StageShared* d_table = nullptr; // [window][4]
    StageShared* h_table = nullptr; // pinned staging of the table
    int* h_fail = nullptr;          // pinned copy of the guard flags
    int window = 256;               // steps per table window = trajectory-ring slots
    std::vector<Bucket> buckets;
    std::vector<std::vector<int>> lists;
    std::vector<std::vector<int64_t>> prefixes;
    std::vector<std::vector<int64_t>> hprefixes;  // pair-symmetric product pass: tiles t0 <= t1
    std::vector<int> tab_list;         // correlated trajectories
    std::vector<int64_t> tab_prefix;   // prefix of their table blocks
    BatchDev b{};
    hipStream_t stream = nullptr;
    Timing timing;
    bool graphs = true;
    bool autotune = true;
    bool tuned = false;
    bool tune_verbose = false;  // G12KBA_TUNE_VERBOSE=1: the autotuner prints every candidate's time
    // pair-symmetric path (k_sym_update): on while every correlated state is
    // pair-symmetric and Hermitian with G1 anti-Hermitian (G2 = 0 is) and every
    // tile size is at most 64; G12KBA_PAIR_SYMMETRY=0 keeps the general path
    bool sym = false;
    bool sym_allowed = true;
    std::vector<char> sym_ok;          // per trajectory: its uploaded state qualifies
    std::vector<int64_t> sym_prefix;   // update blocks before each correlated trajectory
    // unit tables of the update pass (su_table), one per correlated N <= 64
    std::vector<int> su_N;
    std::vector<std::vector<uint64_t>> su_units;
    std::vector<std::vector<uint32_t>> su_firsts;
    std::vector<int64_t> su_nblk;
    std::vector<uint64_t*> su_dunit;
    std::vector<uint32_t*> su_dfirst;
    int* d_flag = nullptr;             // device flag of the symmetry check
    hipGraphExec_t gx[2] = {nullptr, nullptr};  // one step: 0 adiabatic switching, 1 production
    void* arena = nullptr;
    int64_t bytes = 0;
    double t = 0.0;     // production time (shared by all trajectories)
    int64_t steps = 0;  // production steps begun
    int64_t jext_steps = -1;  // closure runs: steps the given collision term covers (-1: none given)
};

}  // namespace g12

using g12::Run;

struct g12_ctx {
    Run r;
};

struct g12_batch {
    Run r;
};

namespace {

// Assigns every device pointer of the run from `base` (nullptr: measure only).
int64_t layout(Run& r, char* base) {
    Carver cv{base};
    for (int i = 0; i < r.ntraj; ++i) {
        DevState& d = r.ds[i];
        const bool corr = r.tp[i].corr != 0;
        const int64_t n4 = corr ? d.n4 : 1, n3 = corr ? d.n3 : 1;
        d.S1 = cv.take<double2>(d.n2);
        d.K1 = cv.take<double2>(d.n2);
        d.A1 = cv.take<double2>(d.n2);
        d.Y1 = cv.take<double2>(d.n2);
        d.h = cv.take<double2>(d.n2);
        d.I = cv.take<double2>(d.n2);
        d.G1t = cv.take<double2>(d.n2);
        d.Wxy = cv.take<double>(d.n2);
        d.Wxx = cv.take<double>(d.n2);
        // mixed-precision runs keep Y2 and R2 as complex64 in the blocked
        // layout (N^2 Nq^2 elements, Nq = N rounded up to 4; see kernels.hip)
        const int64_t nq = (d.N + 3) & ~3;
        const int64_t n4c = r.precision != G12_PREC_FP64 ? (corr ? ((int64_t)d.N * d.N * nq * nq + 1) / 2 : 1) : n4;
        for (int s = 0; s < 2; ++s) {
            d.S2[s] = cv.take<double2>(d.n4);  // the state is kept even when frozen
            d.Y2[s] = cv.take<double2>(n4c);
            d.Sc[s] = r.precision != G12_PREC_FP64 ? cv.take<double2>(n4c) : nullptr;
            // mixed runs: also two blocked complex64 arrays (the stage terms K1, K2
            // of the pair-symmetric update), N^2 Nq^2 complex128 slots
            d.A2[s] = cv.take<double2>(r.precision != G12_PREC_FP64 && corr ? std::max(n4, 2 * n4c) : n4);
            d.R2[s] = cv.take<double2>(n4c);
            d.D[s] = cv.take<double2>(n3);
            d.F[s] = cv.take<double2>(n3);
            d.Qp[s] = cv.take<double2>(n3);
        }
        d.su_unit = nullptr;
        d.su_first = nullptr;
        d.su_nblk = 0;
        d.traj_slots = r.window;
        d.traj = cv.take<double2>(d.n2 * r.window);
        d.jtraj = corr ? cv.take<double2>(d.n2 * r.window) : nullptr;
        d.jext = corr ? nullptr : cv.take<double2>(d.n2 * (r.window + 1));
        d.jext_on = corr ? nullptr : cv.take<int>(1);
    }
    r.d_ds = cv.take<DevState>(r.ntraj);
    r.d_tp = cv.take<TrajParams>(r.ntraj);
    r.d_fail = cv.take<int>(r.ntraj);
    for (int i = 0; i < r.ntraj; ++i) r.ds[i].fail = r.d_fail ? r.d_fail + i : nullptr;
    r.d_counter = cv.take<int>(1);
    r.d_table = cv.take<StageShared>(4 * (int64_t)r.window);
    for (size_t k = 0; k < r.buckets.size(); ++k) {
        r.buckets[k].list = cv.take<int>(r.buckets[k].n);
        r.buckets[k].prefix = cv.take<int64_t>(r.buckets[k].n + 1);
        r.buckets[k].hprefix = cv.take<int64_t>(r.buckets[k].n + 1);
    }
    r.b.tlist = cv.take<int>((int64_t)r.tab_list.size());
    r.b.tprefix = cv.take<int64_t>((int64_t)r.tab_prefix.size());
    r.b.sprefix = cv.take<int64_t>((int64_t)r.sym_prefix.size());
    r.su_dunit.assign(r.su_N.size(), nullptr);
    r.su_dfirst.assign(r.su_N.size(), nullptr);
    for (size_t k = 0; k < r.su_N.size(); ++k) {
        r.su_dunit[k] = cv.take<uint64_t>((int64_t)r.su_units[k].size());
        r.su_dfirst[k] = cv.take<uint32_t>((int64_t)r.su_firsts[k].size());
        for (int i = 0; i < r.ntraj; ++i)
            if (r.tp[i].corr && r.ds[i].N == r.su_N[k]) {
                r.ds[i].su_unit = r.su_dunit[k];
                r.ds[i].su_first = r.su_dfirst[k];
                r.ds[i].su_nblk = (int)r.su_nblk[k];
            }
    }
    r.d_flag = cv.take<int>(1);
    return cv.off;
}

void destroy_run(Run& r) {
    if (r.stream) hipStreamSynchronize(r.stream);
    r.timing.destroy();
    for (auto& g : r.gx)
        if (g) hipGraphExecDestroy(g);
    if (r.arena) hipFree(r.arena);
    if (r.h_table) hipHostFree(r.h_table);
    if (r.h_fail) hipHostFree(r.h_fail);
    if (r.stream) hipStreamDestroy(r.stream);
}

void drop_graphs(Run& r);

// Chooses the pair-symmetric path when every correlated trajectory's state
// qualifies and every tile size is at most SYM_MAX_N; a change of path drops
// the captured steps and the tuned launch shapes.
void update_sym(Run& r) {
    bool on = r.sym_allowed && r.any_corr;
    for (const Bucket& k : r.buckets) on = on && k.NP <= g12::SYM_MAX_N;
    for (int i = 0; i < r.ntraj && on; ++i) on = !r.tp[i].corr || r.sym_ok[i];
    if (on == r.sym) return;
    r.sym = on;
    drop_graphs(r);
    r.tuned = false;
}

int create_run(Run& r, const g12_params* p, int ntraj) {
    if (!p || ntraj < 1) return G12_E_ARG;
    for (int i = 0; i < ntraj; ++i) {
        if (!valid_geometry(p[i].Ns, p[i].Nb)) return G12_E_SIZE;
        if (!valid_precision_size(p[i].Ns * p[i].Nb, p[i].precision)) {
            std::fprintf(stderr, "g12kba: trajectory %d: N = %d exceeds %d, the mixed-precision limit (use FP64)\n", i,
                         p[i].Ns * p[i].Nb, G12_MAX_N_MIXED);
            return G12_E_SIZE;
        }
        if (!same_protocol(p[0], p[i])) return G12_E_ARG;
        if (!valid_protocol(p[i])) {
            std::fprintf(stderr,
                         "g12kba: trajectory %d: a non-finite parameter, AS_rate < 0, AS_dt <= 0 or a time step "
                         "that never reaches time_steps * dt\n",
                         i);
            return G12_E_ARG;
        }
        if (as_step_count(p[i]) < 0) {
            std::fprintf(stderr, "g12kba: trajectory %d: adiabatic switching does not end within %d steps\n", i,
                         MAX_AS_STEPS);
            return G12_E_ARG;
        }
    }
    if (select_device() < 0) return G12_E_HIP;
    g12::device_cu_count();  // cached before any graph capture
    r.ntraj = ntraj;
    r.p.assign(p, p + ntraj);
    r.precision = p[0].precision;
    r.ds.assign(ntraj, DevState{});
    r.tp.assign(ntraj, TrajParams{});
    r.as_steps.assign(ntraj, 0);
    for (int i = 0; i < ntraj; ++i) {
        const g12_params& q = p[i];
        DevState& d = r.ds[i];
        d.Ns = q.Ns;
        d.Nb = q.Nb;
        d.N = q.Ns * q.Nb;
        d.n2 = (int64_t)d.N * d.N;
        d.n3 = d.n2 * d.N;
        d.n4 = d.n2 * d.n2;
        TrajParams& t = r.tp[i];
        t.Ns = q.Ns;
        t.Nb = q.Nb;
        t.N = d.N;
        t.corr = correlated(q);
        t.Nq = q.Nq;
        t.U = q.U;
        t.t1 = q.t1;
        t.t2 = q.t2;
        t.qstr = q.quench_strength;
        t.nV = v_profile(q, t.Vprof);
        for (int o = 0; o < G12_MAX_NB; ++o) t.eps[o] = o < q.Nb ? q.eps[o] : 0.0;
        r.as_steps[i] = as_step_count(q);  // finite: checked above
        t.as_steps = r.as_steps[i];
        r.max_as = std::max(r.max_as, t.as_steps);
        r.maxN = std::max(r.maxN, d.N);
        r.any_corr = r.any_corr || t.corr;
    }
    // buckets of correlated trajectories by padded tile size, larger N first
    // (NP > 64: the wide-tile kernels)
    for (int NP = 16; NP <= G12_MAX_N; NP += 16) {
        std::vector<int> l;
        for (int i = 0; i < ntraj; ++i)
            if (r.tp[i].corr && (r.tp[i].N + 15) / 16 * 16 == NP) l.push_back(i);
        if (l.empty()) continue;
        std::stable_sort(l.begin(), l.end(), [&](int a, int b) { return r.tp[a].N > r.tp[b].N; });
        std::vector<int64_t> pre(l.size() + 1, 0), hpre(l.size() + 1, 0);
        for (size_t j = 0; j < l.size(); ++j) {
            const int64_t N = r.ds[l[j]].N;
            pre[j + 1] = pre[j] + 2 * r.ds[l[j]].n2;
            hpre[j + 1] = hpre[j] + 2 * 16 * g12::hgroups((int)N);
        }
        Bucket k{};
        k.NP = NP;
        k.n = (int)l.size();
        k.tiles = pre.back();
        k.htiles = hpre.back();
        k.full = r.tp[l.back()].N == NP ? 1 : 0;  // l is sorted by N, largest first
        k.waves = k.waves_fused = g12::default_waves(NP);
        r.buckets.push_back(k);
        r.lists.push_back(l);
        r.prefixes.push_back(pre);
        r.hprefixes.push_back(hpre);
    }
    // table launches: N x table_blocks_per_x(N) blocks per correlated trajectory
    r.tab_prefix.assign(1, 0);
    for (int i = 0; i < ntraj; ++i)
        if (r.tp[i].corr) {
            const int N = r.ds[i].N;
            r.tab_list.push_back(i);
            r.tab_prefix.push_back(r.tab_prefix.back() + (int64_t)N * g12::table_blocks_per_x(N));
        }
    // pair-symmetric update: one unit table per correlated N (N <= 64), the
    // same for both spins of every trajectory of that N
    for (int i : r.tab_list) {
        const int N = r.ds[i].N;
        if (N > g12::SYM_MAX_N || std::find(r.su_N.begin(), r.su_N.end(), N) != r.su_N.end()) continue;
        r.su_N.push_back(N);
        r.su_units.emplace_back();
        r.su_firsts.emplace_back();
        r.su_nblk.push_back(g12::su_table(N, r.su_units.back(), r.su_firsts.back()));
    }
    r.sym_prefix.assign(1, 0);
    for (int i : r.tab_list) {
        const int N = r.ds[i].N;
        const auto it = std::find(r.su_N.begin(), r.su_N.end(), N);
        const int64_t nb = it == r.su_N.end() ? 0 : r.su_nblk[it - r.su_N.begin()];
        r.sym_prefix.push_back(r.sym_prefix.back() + 2 * nb);  // a block never holds units of both spins
    }
    const int64_t total = layout(r, nullptr);
    if (hipMalloc(&r.arena, total) != hipSuccess) {
        r.arena = nullptr;
        return G12_E_NOMEM;
    }
    r.bytes = total;
    layout(r, static_cast<char*>(r.arena));
    if (hipHostMalloc((void**)&r.h_table, sizeof(StageShared) * 4 * r.window, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&r.h_fail, sizeof(int) * ntraj, hipHostMallocDefault) != hipSuccess)
        return G12_E_NOMEM;
    HIPCHK(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking));
    // every buffer starts defined (zero): the autotuner runs passes before any
    // stage has written the tables, R2 or Y2
    HIPCHK(hipMemsetAsync(r.arena, 0, total, r.stream));
    HIPCHK(hipMemcpyAsync(r.d_ds, r.ds.data(), sizeof(DevState) * ntraj, hipMemcpyHostToDevice, r.stream));
    HIPCHK(hipMemcpyAsync(r.d_tp, r.tp.data(), sizeof(TrajParams) * ntraj, hipMemcpyHostToDevice, r.stream));
    for (size_t k = 0; k < r.buckets.size(); ++k) {
        HIPCHK(hipMemcpyAsync(const_cast<int*>(r.buckets[k].list), r.lists[k].data(), sizeof(int) * r.lists[k].size(),
                              hipMemcpyHostToDevice, r.stream));
        HIPCHK(hipMemcpyAsync(const_cast<int64_t*>(r.buckets[k].prefix), r.prefixes[k].data(),
                              sizeof(int64_t) * r.prefixes[k].size(), hipMemcpyHostToDevice, r.stream));
        HIPCHK(hipMemcpyAsync(const_cast<int64_t*>(r.buckets[k].hprefix), r.hprefixes[k].data(),
                              sizeof(int64_t) * r.hprefixes[k].size(), hipMemcpyHostToDevice, r.stream));
    }
    if (!r.tab_list.empty()) {
        HIPCHK(hipMemcpyAsync(const_cast<int*>(r.b.tlist), r.tab_list.data(), sizeof(int) * r.tab_list.size(),
                              hipMemcpyHostToDevice, r.stream));
        HIPCHK(hipMemcpyAsync(const_cast<int64_t*>(r.b.tprefix), r.tab_prefix.data(),
                              sizeof(int64_t) * r.tab_prefix.size(), hipMemcpyHostToDevice, r.stream));
        HIPCHK(hipMemcpyAsync(const_cast<int64_t*>(r.b.sprefix), r.sym_prefix.data(),
                              sizeof(int64_t) * r.sym_prefix.size(), hipMemcpyHostToDevice, r.stream));
    }
    for (size_t k = 0; k < r.su_N.size(); ++k) {
        HIPCHK(hipMemcpyAsync(r.su_dunit[k], r.su_units[k].data(), sizeof(uint64_t) * r.su_units[k].size(),
                              hipMemcpyHostToDevice, r.stream));
        HIPCHK(hipMemcpyAsync(r.su_dfirst[k], r.su_firsts[k].data(), sizeof(uint32_t) * r.su_firsts[k].size(),
                              hipMemcpyHostToDevice, r.stream));
    }
    HIPCHK(hipStreamSynchronize(r.stream));
    r.b.ds = r.d_ds;
    r.b.tp = r.d_tp;
    r.b.ntraj = ntraj;
    r.b.maxN = r.maxN;
    r.b.ncorr = (int)r.tab_list.size();
    r.b.tblocks = r.tab_prefix.back();
    r.b.sunits = r.sym_prefix.back();
    r.b.sym_w8 = 1;
    for (int i : r.tab_list) r.b.sym_w8 = r.b.sym_w8 && r.ds[i].N % 16 == 0;
    const char* e = std::getenv("G12KBA_GRAPHS");
    if (e && e[0] == '0') r.graphs = false;
    e = std::getenv("G12KBA_AUTOTUNE");
    if (e && e[0] == '0') r.autotune = false;
    e = std::getenv("G12KBA_TUNE_VERBOSE");
    if (e && e[0] == '1') r.tune_verbose = true;
    e = std::getenv("G12KBA_PAIR_SYMMETRY");
    if (e && e[0] == '0') r.sym_allowed = false;
    r.sym_ok.assign(ntraj, 1);  // every state starts at G1 = G2 = 0
    update_sym(r);
    return 0;
}

template <class F>
void timed(Run& r, int cls, F&& launch) {
    if (!r.timing.on) {
        launch();
        return;
    }
    hipEvent_t a = r.timing.get(), b = r.timing.get();
    hipEventRecord(a, r.stream);
    launch();
    hipEventRecord(b, r.stream);
    r.timing.pending[cls].push_back({a, b});
    if (r.timing.pending[cls].size() > 4096) r.timing.resolve();
}

// One RK4 stage of every trajectory (RK_solver.h:120-203 / AS.h:111-200):
// G1 stage set-up together with the Q/F tables and collision integral,
// strided commutator pass with the sources that are vectors over p1 and q1
// (Psi W and Q terms, Pi^+), G1 right-hand side and update, fused commutator
// pass with the remaining sources (Psi Qp term, Pi) and the G2 update.
// Pair-symmetric path: one product pass over the contiguous tiles (half of the
// commutator and of the sources), then the update pass that adds the
// pair-exchanged half (k_sym_update); timing classes 0 and 1 are these two.
void run_stage(Run& r, const StageRef& sr) {
    hipStream_t s = r.stream;
    // G1 set-up and the tables (from the stage G1 and the partial traces D of
    // the stage input; both tile passes add their share of the sources from them)
    timed(r, 2, [&] { g12::launch_b_stage_setup(r.b, sr, r.precision, s); });
    if (r.sym) {
        for (const Bucket& k : r.buckets)
            timed(r, 0, [&] { g12::launch_b_tile(r.b, k, sr, 0, r.precision, s, 1); });
        timed(r, 2, [&] { g12::launch_b_g1_rhs(r.b, sr, s); });
        timed(r, 1, [&] { g12::launch_b_sym_update(r.b, sr, r.precision, s); });
        return;
    }
    for (const Bucket& k : r.buckets) timed(r, 0, [&] { g12::launch_b_tile(r.b, k, sr, 0, r.precision, s); });
    timed(r, 2, [&] { g12::launch_b_g1_rhs(r.b, sr, s); });
    for (const Bucket& k : r.buckets) timed(r, 1, [&] { g12::launch_b_tile(r.b, k, sr, 1, r.precision, s); });
}

// One whole step, reading its stage blocks at d_table[*d_counter] and
// advancing the counter; production steps begin with the record + guard.
void enqueue_step(Run& r, int phase) {
    StageRef sr{};
    sr.base = r.d_table;
    sr.counter = r.d_counter;
    if (phase == 1) timed(r, 2, [&] { g12::launch_b_record(r.b, sr, r.stream); });
    for (int k = 0; k < 4; ++k) {
        sr.stage = k;
        run_stage(r, sr);
    }
    timed(r, 2, [&] { g12::launch_b_advance_counter(r.d_counter, r.stream); });
}

int capture(Run& r, int phase) {
    if (r.gx[phase]) return 0;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(r.stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(r, phase);
    HIPCHK(hipStreamEndCapture(r.stream, &g));
    const hipError_t e = hipGraphInstantiate(&r.gx[phase], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        r.gx[phase] = nullptr;
        std::fprintf(stderr, "g12kba: hipGraphInstantiate failed: %s\n", hipGetErrorString(e));
        return G12_E_HIP;
    }
    return 0;
}

void drop_graphs(Run& r) {
    for (auto& g : r.gx)
        if (g) {
            hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// Wave counts per bucket and pass, chosen by timing the compiled candidates on
// the run's own tile lists.  Both passes are run as stage 0 of a step (the
// strided pass writes only R2, the fused pass only Y2, A2 and D; the next
// stage 0 overwrites R2 and A2 before reading them, and every run starts, after
// tuning, by rebuilding D and the compact stage-0 input Sc from S), so tuning
// at a step boundary leaves the state unchanged.  Results are cached per
// process by (tile size, precision, tile count).
int tune(Run& r) {
    if (r.tuned || !r.autotune) return 0;
    // (waves, workgroups-per-CU cap) of the strided and the fused pass
    static std::map<std::tuple<int, int, int64_t, int>, std::array<int, 4>> cache;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    StageRef sr{};
    sr.val = stage_block(0, 0.0);
    int rc = 0;
    for (Bucket& k : r.buckets) {
        const auto key = std::make_tuple(k.NP, r.precision, k.tiles, r.sym ? 1 : 0);
        auto it = cache.find(key);
        if (it != cache.end()) {
            k.waves = it->second[0];
            k.wg_per_cu = it->second[1];
            k.waves_fused = it->second[2];
            k.wg_per_cu_fused = it->second[3];
            continue;
        }
        int opts[8];
        const int nopt = g12::tile_wave_options(k.NP, opts);
        // fewer resident workgroups give each more tiles to overlap loads with
        static const int caps[3] = {0, 2, 1};
        std::array<int, 4> best = {k.waves, k.wg_per_cu, k.waves_fused, k.wg_per_cu_fused};
        // the pair-symmetric path has one tile pass (unfused, contiguous tiles)
        for (int fused = 0; fused < (r.sym ? 1 : 2) && !rc; ++fused) {
            float best_ms = 1e30f;
            for (int o = 0; o < nopt && !rc; ++o)
                for (int cap : caps) {
                    Bucket c = k;
                    c.waves = c.waves_fused = opts[o];
                    c.wg_per_cu = c.wg_per_cu_fused = cap;
                    g12::launch_b_tile(r.b, c, sr, fused, r.precision, r.stream, r.sym);  // warm-up
                    hipEventRecord(e0, r.stream);
                    for (int rep = 0; rep < 2; ++rep)
                        g12::launch_b_tile(r.b, c, sr, fused, r.precision, r.stream, r.sym);
                    hipEventRecord(e1, r.stream);
                    if (hipEventSynchronize(e1) != hipSuccess) {
                        rc = G12_E_HIP;
                        break;
                    }
                    float ms = 0.f;
                    hipEventElapsedTime(&ms, e0, e1);
                    if (r.tune_verbose)
                        std::fprintf(stderr, "g12kba tune: NP=%d %s waves=%d wg_cap=%d %.4f ms\n", k.NP,
                                     r.sym ? "products" : fused ? "fused" : "strided", opts[o], cap, ms / 2);
                    if (ms < best_ms) {
                        best_ms = ms;
                        best[2 * fused] = opts[o];
                        best[2 * fused + 1] = cap;
                    }
                }
        }
        k.waves = best[0];
        k.wg_per_cu = best[1];
        k.waves_fused = best[2];
        k.wg_per_cu_fused = best[3];
        cache[key] = best;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    r.tuned = true;
    drop_graphs(r);  // captured launch shapes are stale
    return hipGetLastError() == hipSuccess ? 0 : G12_E_HIP;
}

// Runs `nsteps` steps of a phase in windows of r.window steps.  fill(step,
// slot, blocks) writes the four stage blocks of a step; after(first, n) runs
// once a window has completed (stream synchronised) and returns < 0 on error,
// 1 to stop, 0 to go on.
template <class Fill, class After>
int run_windows(Run& r, int phase, int64_t nsteps, Fill&& fill, After&& after) {
    int rc = tune(r);
    if (rc) return rc;
    const bool graph = r.graphs && !r.timing.on;
    if (graph && (rc = capture(r, phase))) return rc;
    // partial traces of the first stage's input (S), and its complex64 copy
    // in mixed-precision runs; every fused pass then writes both for the
    // next stage input
    for (const Bucket& k : r.buckets) g12::launch_b_d_from_state(r.b, k, r.precision, r.stream, r.sym ? 1 : 0);
    for (int64_t w0 = 0; w0 < nsteps; w0 += r.window) {
        const int n = (int)std::min<int64_t>(r.window, nsteps - w0);
        for (int i = 0; i < n; ++i) fill(w0 + i, i, r.h_table + 4 * i);
        HIPCHK(hipMemcpyAsync(r.d_table, r.h_table, sizeof(StageShared) * 4 * n, hipMemcpyHostToDevice, r.stream));
        HIPCHK(hipMemsetAsync(r.d_counter, 0, sizeof(int), r.stream));
        for (int i = 0; i < n; ++i) {
            if (graph)
                HIPCHK(hipGraphLaunch(r.gx[phase], r.stream));
            else
                enqueue_step(r, phase);
        }
        HIPCHK(hipMemcpyAsync(r.h_fail, r.d_fail, sizeof(int) * r.ntraj, hipMemcpyDeviceToHost, r.stream));
        HIPCHK(hipStreamSynchronize(r.stream));
        HIPCHK(hipGetLastError());
        rc = after(w0, n);
        if (rc) return rc < 0 ? rc : 0;
    }
    return 0;
}

// Adiabatic switching of every trajectory (AS.h:63-214): fixed step AS_dt,
// U_AS = fermi(t) U per stage with V scaled alike, no quench; trajectory i
// stops after its own as_steps[i] steps.
int run_as(Run& r, int* steps) {
    const g12_params& p = r.p[0];
    double t = 0.0;
    int rc = run_windows(
        r, 0, r.max_as,
        [&](int64_t n, int, StageShared* blk) {
            for (int k = 0; k < 4; ++k) {
                const double dtk = (k == 1 || k == 2) ? p.AS_dt / 2 : p.AS_dt;
                const double ts = k == 0 ? t : t + dtk;
                blk[k] = stage_block(k, p.AS_dt);
                blk[k].lam = g12_fermi(ts, p.AS_midpoint, p.AS_rate);
                blk[k].as_phase = 1;
                blk[k].step = (int)n;
            }
            t += p.AS_dt;
        },
        [&](int64_t, int) { return 0; });
    r.timing.resolve();
    if (rc) return rc;
    if (steps)
        for (int i = 0; i < r.ntraj; ++i) steps[i] = r.as_steps[i];
    return 0;
}

// Quench ramp at stage time ts (RK_solver.h:141-151): 0 before the gate,
// else the switch-on logistic (times the switch-off one for a pulse); the
// shift on sites < Nq is phi * quench_strength.
double quench_phi(const g12_params& p, double ts, int* gate) {
    *gate = 0;
    if (ts < p.quench_gate_time) return 0.0;
    *gate = 1;
    const double on = g12_fermi(ts, p.quench_on_midpoint, p.quench_rate);
    return p.q_type == G12_QUENCH_PULSE ? on * (1 - g12_fermi(ts, p.quench_off_midpoint, p.quench_rate)) : on;
}

// Stage blocks of production step n at time t (RK_solver.h:120-151).
void production_blocks(const g12_params& p, double t, int64_t n, int slot, StageShared* blk) {
    for (int k = 0; k < 4; ++k) {
        const double dtk = (k == 1 || k == 2) ? p.dt / 2 : p.dt;
        const double ts = k == 0 ? t : t + dtk;
        blk[k] = stage_block(k, p.dt);
        blk[k].step = (int)n;
        blk[k].slot = slot;
        blk[k].phi = quench_phi(p, ts, &blk[k].gate);
    }
}

// Production RK4 of every trajectory (RK_solver.h:64-226), `nsteps` steps from
// the run's current time.  The G1 matrix before each step goes to traj[i]
// (capacity cap matrices, from step index `rec0`) when traj is given.
int run_production(Run& r, int64_t nsteps, double* const* traj, int64_t cap, int64_t rec0,
                   double* const* jtraj = nullptr) {
    const g12_params& p = r.p[0];
    // a context given a collision term (closure runs) runs at most the steps
    // it covers, once
    if (r.ntraj == 1 && r.jext_steps >= 0) {
        if (nsteps > r.jext_steps) {
            std::fprintf(stderr, "g12kba: %lld steps asked, the given collision term covers %lld\n",
                         (long long)nsteps, (long long)r.jext_steps);
            return G12_E_ARG;
        }
        r.jext_steps = 0;
    }
    const double t0 = r.t;
    double t = t0;
    const int64_t s0 = r.steps;
    int rc = run_windows(
        r, 1, nsteps,
        [&](int64_t i, int slot, StageShared* blk) {
            production_blocks(p, t, s0 + i, slot, blk);
            t += p.dt;
        },
        [&](int64_t w0, int n) -> int {
            bool all_failed = true;
            for (int i = 0; i < r.ntraj; ++i) {
                const int f = r.h_fail[i];
                if (!f) all_failed = false;
                const DevState& d = r.ds[i];
                const bool want_g = traj && traj[i], want_j = jtraj && jtraj[i] && d.jtraj;
                if (!want_g && !want_j) continue;
                // G1 records of this window: up to and including the failing
                // step; collision terms: the steps that ran
                int64_t hi = s0 + w0 + n;
                if (f && f < hi) hi = f;
                const int64_t lo = s0 + w0;
                const int64_t top = std::min(hi, rec0 + cap);
                if (top <= lo || lo < rec0) continue;
                if (want_g)
                    HIPCHK(hipMemcpyAsync(traj[i] + (lo - rec0) * d.n2 * 2, d.traj,
                                          sizeof(double2) * d.n2 * (top - lo), hipMemcpyDeviceToHost, r.stream));
                const int64_t jtop = (f && f - 1 < top) ? f - 1 : top;
                if (want_j && jtop > lo)
                    HIPCHK(hipMemcpyAsync(jtraj[i] + (lo - rec0) * d.n2 * 2, d.jtraj,
                                          sizeof(double2) * d.n2 * (jtop - lo), hipMemcpyDeviceToHost, r.stream));
            }
            HIPCHK(hipStreamSynchronize(r.stream));
            return all_failed ? 1 : 0;
        });
    r.timing.resolve();
    if (rc) return rc;
    // A run of one trajectory whose guard tripped at step f - 1 (fail = f)
    // holds the state before that step: its clock stops there.  The
    // trajectories of a larger run share one clock, which goes on.
    const int f = r.h_fail[0];
    if (r.ntraj == 1 && f > s0) {
        t = t0;
        for (int64_t i = s0; i < f - 1; ++i) t += p.dt;  // the accumulated time of RK_solver.h:216
        r.t = t;
        r.steps = f - 1;
        return 0;
    }
    r.t = t;
    r.steps = s0 + nsteps;
    return 0;
}

// Production iterations begun by a run of one trajectory in a call that
// started at step s0 and asked for n: n, or up to and including the step
// whose guard tripped.
int64_t iterations_begun(const Run& r, int64_t s0, int64_t n) {
    const int f = r.h_fail[0];
    return (f > s0 && f - s0 < n) ? f - s0 : n;
}

int64_t production_step_count(const g12_params& p) {
    const double tmax = double(p.time_steps) * p.dt;
    int64_t n = 0;
    for (double t = 0.0; t < tmax; t += p.dt) ++n;  // RK_solver.h:96,216
    return n;
}

int reset_guard(Run& r) {
    HIPCHK(hipMemsetAsync(r.d_fail, 0, sizeof(int) * r.ntraj, r.stream));
    HIPCHK(hipStreamSynchronize(r.stream));
    for (int i = 0; i < r.ntraj; ++i) r.h_fail[i] = 0;
    return 0;
}

// Full production run: returns 0, or G12_GUARD if any trajectory's guard
// tripped; steps[i] = trajectory length (iterations begun).
int run_rk4(Run& r, double* const* traj, int64_t cap, int64_t* steps) {
    int rc = reset_guard(r);
    if (rc) return rc;
    r.t = 0.0;
    r.steps = 0;
    const int64_t n = production_step_count(r.p[0]);
    rc = run_production(r, n, traj, cap, 0);
    if (rc) return rc;
    HIPCHK(hipMemcpy(r.h_fail, r.d_fail, sizeof(int) * r.ntraj, hipMemcpyDeviceToHost));
    bool guard = false;
    for (int i = 0; i < r.ntraj; ++i) {
        const int f = r.h_fail[i];
        if (f) guard = true;
        if (steps) steps[i] = f ? f : n;
    }
    return guard ? G12_GUARD : 0;
}

int run_advance(Run& r, int64_t nsteps) {
    int rc = run_production(r, nsteps, nullptr, 0, 0);
    if (rc) return rc;
    for (int i = 0; i < r.ntraj; ++i)
        if (r.h_fail[i]) return G12_GUARD;
    return 0;
}

// Moves trajectory j's G2 state between the natural layout (general path) and
// the Hermitian half layout (pair-symmetric path), through its A buffer (at
// least N^4 complex128 elements; a step-internal buffer, free between steps).
void convert_state(Run& r, int j, bool to_half) {
    const DevState& d = r.ds[j];
    for (int s = 0; s < 2; ++s) {
        (void)hipMemcpyAsync(d.A2[s], d.S2[s], sizeof(double2) * d.n4, hipMemcpyDeviceToDevice, r.stream);
        if (to_half)
            g12::launch_nat_to_half(d.A2[s], d.S2[s], d.N, r.stream);
        else
            g12::launch_half_to_nat(d.A2[s], d.S2[s], d.N, r.stream);
    }
}

int set_state(Run& r, int i, const double* G1, const double* G2xx, const double* G2xy) {
    if (i < 0 || i >= r.ntraj || !G1) return G12_E_ARG;
    const DevState& d = r.ds[i];
    const bool was_sym = r.sym;  // the layout of the other trajectories' states
    HIPCHK(hipMemcpyAsync(d.S1, G1, sizeof(double2) * d.n2, hipMemcpyHostToDevice, r.stream));
    const double* g2[2] = {G2xx, G2xy};
    double2* tmp = nullptr;
    if ((G2xx || G2xy) && hipMalloc(&tmp, sizeof(double2) * d.n4) != hipSuccess) {
        (void)hipStreamSynchronize(r.stream);
        return G12_E_NOMEM;
    }
    hipError_t e = hipSuccess;  // first failure; the stream is drained before tmp is freed
    for (int s = 0; s < 2 && e == hipSuccess; ++s) {
        if (!g2[s]) {
            e = hipMemsetAsync(d.S2[s], 0, sizeof(double2) * d.n4, r.stream);
            continue;
        }
        e = hipMemcpyAsync(tmp, g2[s], sizeof(double2) * d.n4, hipMemcpyHostToDevice, r.stream);
        if (e == hipSuccess) g12::launch_ref_to_nat(tmp, d.S2[s], d.Ns, d.Nb, r.stream);
    }
    if (e == hipSuccess) e = hipMemsetAsync(d.fail, 0, sizeof(int), r.stream);
    // does the state qualify for the pair-symmetric path?  G1 anti-Hermitian and
    // each G2 tensor pair-symmetric and Hermitian, to 1e-12 of its largest
    // element (the path then keeps both symmetries exactly)
    int flag = 0;
    if (r.tp[i].corr && e == hipSuccess) {
        e = hipMemsetAsync(r.d_flag, 0, sizeof(int), r.stream);
        for (int sp = 0; sp < 2 && e == hipSuccess; ++sp) {
            if (!g2[sp]) continue;
            double mx = 0.0;
            for (int64_t k = 0; k < 2 * d.n4; ++k) mx = std::max(mx, std::fabs(g2[sp][k]));
            g12::launch_pair_check(d.S2[sp], d.N, 1e-12 * mx, 1, r.d_flag, r.stream);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(&flag, r.d_flag, sizeof(int), hipMemcpyDeviceToHost, r.stream);
    }
    const hipError_t es = hipStreamSynchronize(r.stream);
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess || es != hipSuccess) return G12_E_HIP;
    if (r.tp[i].corr) {
        double mx = 0.0, asym = 0.0;
        for (int64_t k = 0; k < 2 * d.n2; ++k) mx = std::max(mx, std::fabs(G1[k]));
        for (int x = 0; x < d.N; ++x)
            for (int y = 0; y < d.N; ++y) {  // G1 + G1^dagger
                const int64_t xy = 2 * ((int64_t)x * d.N + y), yx = 2 * ((int64_t)y * d.N + x);
                asym = std::max(asym, std::max(std::fabs(G1[xy] + G1[yx]), std::fabs(G1[xy + 1] - G1[yx + 1])));
            }
        r.sym_ok[i] = (flag == 0 && asym <= 1e-12 * mx) ? 1 : 0;
        update_sym(r);
        // trajectory i's state is natural now, the others in the previous
        // path's layout: bring every correlated state to the current one
        for (int j = 0; j < r.ntraj; ++j) {
            if (!r.tp[j].corr) continue;
            const bool half_now = j != i && was_sym;
            if (half_now != r.sym) convert_state(r, j, r.sym);
        }
        HIPCHK(hipStreamSynchronize(r.stream));
    }
    r.h_fail[i] = 0;
    r.t = 0.0;
    r.steps = 0;
    return 0;
}

int get_state(Run& r, int i, double* G1, double* G2xx, double* G2xy) {
    if (i < 0 || i >= r.ntraj) return G12_E_ARG;
    const DevState& d = r.ds[i];
    if (G1) HIPCHK(hipMemcpyAsync(G1, d.S1, sizeof(double2) * d.n2, hipMemcpyDeviceToHost, r.stream));
    double* g2[2] = {G2xx, G2xy};
    const bool half = r.sym && r.tp[i].corr;  // the state is in the Hermitian half layout
    double2* tmp = nullptr;
    if ((G2xx || G2xy) && hipMalloc(&tmp, sizeof(double2) * d.n4 * (half ? 2 : 1)) != hipSuccess) {
        (void)hipStreamSynchronize(r.stream);
        return G12_E_NOMEM;
    }
    hipError_t e = hipSuccess;  // first failure; the stream is drained before tmp is freed
    for (int s = 0; s < 2 && e == hipSuccess; ++s) {
        if (!g2[s]) continue;
        const double2* nat = d.S2[s];
        if (half) {
            g12::launch_half_to_nat(d.S2[s], tmp + d.n4, d.N, r.stream);
            nat = tmp + d.n4;
        }
        g12::launch_nat_to_ref(nat, tmp, d.Ns, d.Nb, r.stream);
        e = hipMemcpyAsync(g2[s], tmp, sizeof(double2) * d.n4, hipMemcpyDeviceToHost, r.stream);
    }
    const hipError_t es = hipStreamSynchronize(r.stream);
    if (tmp) (void)hipFree(tmp);
    return (e == hipSuccess && es == hipSuccess) ? 0 : G12_E_HIP;
}

int timing_enable(Run& r, int on) {
    r.timing.resolve();
    r.timing.on = on != 0;
    for (int i = 0; i < 3; ++i) {
        r.timing.count[i] = 0;
        r.timing.ms[i] = 0;
    }
    return 0;
}

int timing_get(Run& r, int which, int64_t* launches, double* total_ms) {
    if (which < 0 || which > 2) return G12_E_ARG;
    r.timing.resolve();
    if (launches) *launches = r.timing.count[which];
    if (total_ms) *total_ms = r.timing.ms[which];
    return 0;
}

int configure(Run& r, int use_graphs, int autotune) {
    if (use_graphs >= 0) r.graphs = use_graphs != 0;
    if (autotune >= 0) {
        const bool want = autotune != 0;
        if (!want && r.tuned) {  // back to the compiled defaults
            for (Bucket& k : r.buckets) {
                k.waves = k.waves_fused = g12::default_waves(k.NP);
                k.wg_per_cu = k.wg_per_cu_fused = 0;
            }
            drop_graphs(r);
        }
        r.autotune = want;
        r.tuned = false;
    }
    return 0;
}

}  // namespace

extern "C" {

const char* g12_version(void) { return "g12kba-mi355x 0.2"; }

int g12_create(const g12_params* p, g12_ctx** out) {
    if (!p || !out) return G12_E_ARG;
    if (!valid_geometry(p->Ns, p->Nb)) return G12_E_SIZE;
    g12_ctx* c = new (std::nothrow) g12_ctx();
    if (!c) return G12_E_NOMEM;
    const int rc = create_run(c->r, p, 1);
    if (rc) {
        destroy_run(c->r);
        delete c;
        return rc;
    }
    *out = c;
    return 0;
}

int g12_destroy(g12_ctx* c) {
    if (!c) return 0;
    destroy_run(c->r);
    delete c;
    return 0;
}

int g12_synchronize(void) {
    if (select_device() < 0) return G12_E_HIP;
    return hipDeviceSynchronize() == hipSuccess ? 0 : G12_E_HIP;
}

int g12_memory_bytes(g12_ctx* c, int64_t* bytes) {
    if (!c || !bytes) return G12_E_ARG;
    *bytes = c->r.bytes;
    return 0;
}

int g12_set_state(g12_ctx* c, const double* G1, const double* G2xx, const double* G2xy) {
    if (!c) return G12_E_ARG;
    return set_state(c->r, 0, G1, G2xx, G2xy);
}

int g12_get_state(g12_ctx* c, double* G1, double* G2xx, double* G2xy) {
    if (!c) return G12_E_ARG;
    return get_state(c->r, 0, G1, G2xx, G2xy);
}

int g12_get_time(g12_ctx* c, double* t, int64_t* steps) {
    if (!c) return G12_E_ARG;
    if (t) *t = c->r.t;
    if (steps) *steps = c->r.steps;
    return 0;
}

int g12_adiabatic_switching(g12_ctx* c, int* steps) {
    if (!c) return G12_E_ARG;
    return run_as(c->r, steps);
}

int g12_rk4(g12_ctx* c, double* traj, int64_t traj_cap, int64_t* steps) {
    if (!c) return G12_E_ARG;
    double* tr[1] = {traj};
    return run_rk4(c->r, traj ? tr : nullptr, traj_cap, steps);
}

int g12_rk4_advance(g12_ctx* c, int64_t nsteps, int64_t* done) {
    if (!c || nsteps < 0) return G12_E_ARG;
    if (c->r.h_fail[0]) {  // stopped by the guard: nothing runs until set_state / load_state
        if (done) *done = 0;
        return G12_GUARD;
    }
    const int64_t s0 = c->r.steps;
    const int rc = run_advance(c->r, nsteps);
    if (rc < 0) return rc;
    if (done) *done = iterations_begun(c->r, s0, nsteps);
    return rc;
}

int g12_rk4_record(g12_ctx* c, int64_t nsteps, double* traj, int64_t traj_cap, int64_t* done) {
    if (!c || nsteps < 0) return G12_E_ARG;
    Run& r = c->r;
    if (r.h_fail[0]) {  // stopped by the guard: nothing runs until set_state / load_state
        if (done) *done = 0;
        return G12_GUARD;
    }
    double* tr[1] = {traj};
    const int64_t s0 = r.steps;
    const int rc = run_production(r, nsteps, traj ? tr : nullptr, traj_cap, s0);
    if (rc) return rc;
    if (done) *done = iterations_begun(r, s0, nsteps);
    return r.h_fail[0] ? G12_GUARD : 0;
}

int g12_rk4_record_collision(g12_ctx* c, int64_t nsteps, double* traj, double* jtraj, int64_t traj_cap,
                              int64_t* done) {
    if (!c || nsteps < 0) return G12_E_ARG;
    Run& r = c->r;
    if (!r.tp[0].corr && jtraj) return G12_E_ARG;  // no collision term to record
    if (r.h_fail[0]) {
        if (done) *done = 0;
        return G12_GUARD;
    }
    double* tr[1] = {traj};
    double* jt[1] = {jtraj};
    const int64_t s0 = r.steps;
    const int rc = run_production(r, nsteps, traj ? tr : nullptr, traj_cap, s0, jtraj ? jt : nullptr);
    if (rc) return rc;
    if (done) *done = iterations_begun(r, s0, nsteps);
    return r.h_fail[0] ? G12_GUARD : 0;
}

int g12_set_collision(g12_ctx* c, const double* J, int64_t count) {
    if (!c) return G12_E_ARG;
    Run& r = c->r;
    const DevState& d = r.ds[0];
    if (r.tp[0].corr || !d.jext) return G12_E_ARG;  // only uncorrelated (HF) contexts take a given term
    const int on = (J && count >= 2) ? 1 : 0;
    if (J && (count < 2 || count > r.window + 1)) return G12_E_ARG;
    if (on) HIPCHK(hipMemcpyAsync(d.jext, J, sizeof(double2) * d.n2 * count, hipMemcpyHostToDevice, r.stream));
    HIPCHK(hipMemcpyAsync(d.jext_on, &on, sizeof(int), hipMemcpyHostToDevice, r.stream));
    HIPCHK(hipStreamSynchronize(r.stream));
    r.jext_steps = on ? count - 1 : -1;
    return 0;
}

int g12_set_time(g12_ctx* c, double t, int64_t steps) {
    if (!c || !std::isfinite(t) || steps < 0) return G12_E_ARG;
    c->r.t = t;
    c->r.steps = steps;
    return 0;
}

int g12_timing_enable(g12_ctx* c, int on) { return c ? timing_enable(c->r, on) : G12_E_ARG; }

int g12_timing_get(g12_ctx* c, int which, int64_t* launches, double* total_ms) {
    return c ? timing_get(c->r, which, launches, total_ms) : G12_E_ARG;
}

int g12_pair_symmetric(g12_ctx* c, int* on) {
    if (!c || !on) return G12_E_ARG;
    *on = c->r.sym ? 1 : 0;
    return 0;
}

int g12_configure(g12_ctx* c, int use_graphs, int autotune) {
    return c ? configure(c->r, use_graphs, autotune) : G12_E_ARG;
}

// ---- runs of many trajectories ---------------------------------------------

int g12_batch_create(const g12_params* p, int ntraj, g12_batch** out) {
    if (!p || !out || ntraj < 1) return G12_E_ARG;
    g12_batch* b = new (std::nothrow) g12_batch();
    if (!b) return G12_E_NOMEM;
    const int rc = create_run(b->r, p, ntraj);
    if (rc) {
        destroy_run(b->r);
        delete b;
        return rc;
    }
    *out = b;
    return 0;
}

int g12_batch_destroy(g12_batch* b) {
    if (!b) return 0;
    destroy_run(b->r);
    delete b;
    return 0;
}

int g12_batch_size(g12_batch* b, int* ntraj) {
    if (!b || !ntraj) return G12_E_ARG;
    *ntraj = b->r.ntraj;
    return 0;
}

int g12_batch_memory_bytes(g12_batch* b, int64_t* bytes) {
    if (!b || !bytes) return G12_E_ARG;
    *bytes = b->r.bytes;
    return 0;
}

int g12_batch_set_state(g12_batch* b, int i, const double* G1, const double* G2xx, const double* G2xy) {
    return b ? set_state(b->r, i, G1, G2xx, G2xy) : G12_E_ARG;
}

int g12_batch_get_state(g12_batch* b, int i, double* G1, double* G2xx, double* G2xy) {
    return b ? get_state(b->r, i, G1, G2xx, G2xy) : G12_E_ARG;
}

int g12_batch_adiabatic_switching(g12_batch* b, int* steps) { return b ? run_as(b->r, steps) : G12_E_ARG; }

int g12_batch_rk4(g12_batch* b, double* const* traj, int64_t traj_cap, int64_t* steps) {
    return b ? run_rk4(b->r, traj, traj_cap, steps) : G12_E_ARG;
}

int g12_batch_rk4_advance(g12_batch* b, int64_t nsteps) {
    if (!b || nsteps < 0) return G12_E_ARG;
    return run_advance(b->r, nsteps);
}

int g12_batch_rk4_record(g12_batch* b, int64_t nsteps, double* const* traj, int64_t traj_cap) {
    if (!b || nsteps < 0) return G12_E_ARG;
    Run& r = b->r;
    const int rc = run_production(r, nsteps, traj, traj_cap, r.steps);
    if (rc) return rc;
    for (int i = 0; i < r.ntraj; ++i)
        if (r.h_fail[i]) return G12_GUARD;
    return 0;
}

int g12_batch_rk4_record_collision(g12_batch* b, int64_t nsteps, double* const* traj, double* const* jtraj,
                                    int64_t traj_cap) {
    if (!b || nsteps < 0) return G12_E_ARG;
    Run& r = b->r;
    const int rc = run_production(r, nsteps, traj, traj_cap, r.steps, jtraj);
    if (rc) return rc;
    for (int i = 0; i < r.ntraj; ++i)
        if (r.h_fail[i]) return G12_GUARD;
    return 0;
}

int g12_batch_guard(g12_batch* b, int* fail) {
    if (!b || !fail) return G12_E_ARG;
    HIPCHK(hipMemcpy(fail, b->r.d_fail, sizeof(int) * b->r.ntraj, hipMemcpyDeviceToHost));
    return 0;
}

int g12_batch_timing_enable(g12_batch* b, int on) { return b ? timing_enable(b->r, on) : G12_E_ARG; }

int g12_batch_timing_get(g12_batch* b, int which, int64_t* launches, double* total_ms) {
    return b ? timing_get(b->r, which, launches, total_ms) : G12_E_ARG;
}

int g12_batch_configure(g12_batch* b, int use_graphs, int autotune) {
    return b ? configure(b->r, use_graphs, autotune) : G12_E_ARG;
}

int g12_batch_tune(g12_batch* b) { return b ? tune(b->r) : G12_E_ARG; }

int g12_batch_pair_symmetric(g12_batch* b, int* on) {
    if (!b || !on) return G12_E_ARG;
    *on = b->r.sym ? 1 : 0;
    return 0;
}

int g12_batch_buckets(g12_batch* b, int cap, int* NP, int* count, int* waves_strided, int* waves_fused) {
    if (!b) return G12_E_ARG;
    const Run& r = b->r;
    const int n = (int)r.buckets.size();
    for (int k = 0; k < n && k < cap; ++k) {
        if (NP) NP[k] = r.buckets[k].NP;
        if (count) count[k] = r.buckets[k].n;
        if (waves_strided) waves_strided[k] = r.buckets[k].waves;
        if (waves_fused) waves_fused[k] = r.buckets[k].waves_fused;
    }
    return n;
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

// A one-trajectory batch on scratch buffers through which the routine-level
// entries run the production kernels: the stage set-up (G1 part: Y1, its
// conjugate transpose, h_HF and the W tables; Qp and F tables and the
// collision integral) and the tile passes with their source epilogues.  The
// interactions are given as V_n themselves (v_abs), and the stage is stage 0 of
// a production step unless set otherwise (Y1 = S1, stage input S2).
struct RoutineBatch {
    Scratch sc;
    DevState d{};
    TrajParams tp{};
    BatchDev b{};
    Bucket k{};
    hipStream_t s = api_stream();

    // with_g2: allocate the G2 tensors and tables and enable the table roles
    int init(int Ns, int Nb, double U, const double* V, int nV, bool with_g2) {
        const int N = Ns * Nb;
        d.Ns = Ns;
        d.Nb = Nb;
        d.N = N;
        d.n2 = (int64_t)N * N;
        d.n3 = d.n2 * N;
        d.n4 = d.n2 * d.n2;
        std::vector<std::pair<void*, int64_t>> zero;
        auto c2 = [&](int64_t n) {
            double2* p = sc.get<double2>(n);
            zero.push_back({p, (int64_t)sizeof(double2) * n});
            return p;
        };
        d.S1 = c2(d.n2);
        d.K1 = c2(d.n2);
        d.A1 = c2(d.n2);
        d.Y1 = c2(d.n2);
        d.h = c2(d.n2);
        d.I = c2(d.n2);
        d.G1t = c2(d.n2);
        d.Wxy = sc.get<double>(d.n2);
        d.Wxx = sc.get<double>(d.n2);
        if (with_g2)
            for (int sp = 0; sp < 2; ++sp) {
                d.S2[sp] = c2(d.n4);
                d.Y2[sp] = c2(d.n4);
                d.A2[sp] = c2(d.n4);
                d.R2[sp] = c2(d.n4);
                d.D[sp] = c2(d.n3);
                d.F[sp] = c2(d.n3);
                d.Qp[sp] = c2(d.n3);
            }
        for (auto& z : zero)
            if (!z.first) return G12_E_NOMEM;
        if (!d.Wxy || !d.Wxx) return G12_E_NOMEM;
        for (auto& z : zero)
            if (hipMemsetAsync(z.first, 0, z.second, s) != hipSuccess) return G12_E_HIP;
        tp.Ns = Ns;
        tp.Nb = Nb;
        tp.N = N;
        tp.corr = with_g2 ? 1 : 0;
        tp.U = U;
        tp.nV = nV < G12_MAX_SHELLS ? nV : G12_MAX_SHELLS;
        tp.v_abs = 1;
        for (int i = 0; i < tp.nV; ++i) tp.Vprof[i] = V[i];
        return 0;
    }

    // descriptors on the device; call after the last change to d and tp
    int upload() {
        const int N = d.N;
        const int64_t tblocks = tp.corr ? (int64_t)N * g12::table_blocks_per_x(N) : 0;
        DevState* dds = sc.get<DevState>(1);
        TrajParams* dtp = sc.get<TrajParams>(1);
        int* list = sc.get<int>(1);
        int64_t* prefix = sc.get<int64_t>(2);
        int64_t* tprefix = sc.get<int64_t>(2);
        DEV_OR_FAIL(dds && dtp && list && prefix && tprefix);
        const int zero_i = 0;
        const int64_t pre[2] = {0, 2 * d.n2}, tpre[2] = {0, tblocks};
        int rc;
        if ((rc = up(dds, &d, sizeof(DevState))) || (rc = up(dtp, &tp, sizeof(TrajParams))) ||
            (rc = up(list, &zero_i, sizeof(int))) || (rc = up(prefix, pre, sizeof(pre))) ||
            (rc = up(tprefix, tpre, sizeof(tpre))))
            return rc;
        b.ds = dds;
        b.tp = dtp;
        b.ntraj = 1;
        b.maxN = N;
        b.tlist = list;
        b.tprefix = tprefix;
        b.ncorr = tp.corr ? 1 : 0;
        b.tblocks = tblocks;
        k.NP = (N + 15) / 16 * 16;
        k.list = list;
        k.prefix = prefix;
        k.n = 1;
        k.tiles = 2 * d.n2;
        k.waves = k.waves_fused = g12::default_waves(k.NP);
        k.full = N == k.NP ? 1 : 0;
        return 0;
    }

    // stage scalars: Y1 = S1, the next stage input S + cn K, A = w K (first) or A + w K
    static StageRef stage(int first, double wk, double cn) {
        StageRef r{};
        r.val.wk = wk;
        r.val.cn = cn;
        r.val.first = first;
        r.val.lam = 1.0;
        return r;
    }

    int upload_nat(double2* dst, const double* ref, double2* raw) {  // reference layout -> natural
        const int rc = up(raw, ref, sizeof(double2) * d.n4);
        if (!rc) g12::launch_ref_to_nat(raw, dst, d.Ns, d.Nb, s);
        return rc;
    }
    int download_ref(double* ref, const double2* src, double2* raw) {  // natural -> reference layout
        g12::launch_nat_to_ref(src, raw, d.Ns, d.Nb, s);
        return down(ref, raw, sizeof(double2) * d.n4);
    }
};

// x <- i x for n complex values (exact: a swap and a sign)
void times_i(double* x, int64_t n) {
    for (int64_t e = 0; e < n; ++e) {
        const double re = x[2 * e];
        x[2 * e] = -x[2 * e + 1];
        x[2 * e + 1] = re;
    }
}

// ---------------------------------------------------------------------------
// energy.  No reference counterpart (README.md:61-63 lists energies as future
// work).  With rho = -i G1 per spin (the reference's Hartree term -iU G1_XX
// is +U n_X) and the interaction tables W_xx, W_xy scaled by lam:
//   one-body     2 Re sum_XY h0[X][Y] rho[Y][X]   (hopping, eps, quench shift w
//                on sites < Nq: the non-interacting part of h_HFx_quench, HF.h:57-81)
//   Hartree      Re sum_XY (W_xx + W_xy)(X,Y) rho[X][X] rho[Y][Y]
//   exchange     - sum_XY W_xx(X,Y) |rho[X][Y]|^2
//   correlation  - Re sum_XY ( W_xx(X,Y) M^xx[(X,Y);(X,Y)] + W_xy(X,Y) M^xy[(X,Y);(X,Y)] )
// The weight -1 of the correlation term is the one for which the equations of
// motion (G1_G2_t_step.h:24-86) conserve the total: for a time-independent h0
// the total is constant along the exact dynamics when the Fock term is the
// true exchange.  The reference's onsite
// inter-orbital Fock term takes the transposed G1 element (HF.h:83), so for
// Nb >= 2 the total drifts once those elements become complex; for Nb = 1 it
// is conserved to the integrator's error.
template <class Diag>
void energy_parts(const g12_params& p, double lam, double w, const double* G1_, Diag diag, bool corr, double* E) {
    typedef std::complex<double> cplx;
    const cplx* G1 = reinterpret_cast<const cplx*>(G1_);
    const int Ns = p.Ns, Nb = p.Nb, N = Ns * Nb;
    double prof[G12_MAX_SHELLS];
    const int nV = v_profile(p, prof);
    auto wxy = [&](int X, int Y) {
        int d = X % Ns - Y % Ns;
        d = d < 0 ? -d : d;
        if (d == 0) return lam * p.U;
        return d <= nV ? lam * p.U * prof[d - 1] : 0.0;
    };
    auto rho = [&](int X, int Y) { return cplx(0, -1) * G1[(int64_t)X * N + Y]; };
    double e1 = 0, eh = 0, ef = 0, ec = 0;
    for (int X = 0; X < N; ++X) {
        const int sx = X % Ns, ox = X / Ns;
        // h0 row X: eps + quench shift, -t1 to the neighbouring sites, -t2 to the adjacent orbitals
        double diag0 = p.eps[ox] + (sx < p.Nq ? w : 0.0);
        e1 += 2 * (diag0 * rho(X, X)).real();
        if (sx > 0) e1 += 2 * (-p.t1 * rho(X - 1, X)).real();
        if (sx + 1 < Ns) e1 += 2 * (-p.t1 * rho(X + 1, X)).real();
        if (ox > 0) e1 += 2 * (-p.t2 * rho(X - Ns, X)).real();
        if (ox + 1 < Nb) e1 += 2 * (-p.t2 * rho(X + Ns, X)).real();
        for (int Y = 0; Y < N; ++Y) {
            const double a = wxy(X, Y), b = X == Y ? 0.0 : a;
            if (a == 0.0) continue;
            eh += ((a + b) * rho(X, X) * rho(Y, Y)).real();
            ef -= b * std::norm(rho(X, Y));
            if (corr) ec -= (b * diag(0, X, Y) + a * diag(1, X, Y)).real();
        }
    }
    E[0] = e1;
    E[1] = eh;
    E[2] = ef;
    E[3] = ec;
}

int state_energy(Run& r, int i, double* E) {
    if (i < 0 || i >= r.ntraj || !E) return G12_E_ARG;
    const DevState& d = r.ds[i];
    const g12_params& p = r.p[i];
    const bool corr = r.tp[i].corr != 0;
    std::vector<std::complex<double>> g1(d.n2), pd(corr ? 2 * d.n2 : 0);
    double2* tmp = nullptr;
    if (corr && hipMalloc(&tmp, sizeof(double2) * 2 * d.n2) != hipSuccess) return G12_E_NOMEM;
    hipError_t e = hipSuccess;  // first failure; the stream is drained before tmp and the host buffers go
    if (corr) {
        for (int s = 0; s < 2; ++s) g12::launch_pair_diag(d.S2[s], tmp + s * d.n2, d.N, r.stream, r.sym ? 1 : 0);
        e = hipMemcpyAsync(pd.data(), tmp, sizeof(double2) * 2 * d.n2, hipMemcpyDeviceToHost, r.stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(g1.data(), d.S1, sizeof(double2) * d.n2, hipMemcpyDeviceToHost, r.stream);
    const hipError_t es = hipStreamSynchronize(r.stream);
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess || es != hipSuccess) return G12_E_HIP;
    int gate = 0;
    const double w = quench_phi(p, r.t, &gate) * p.quench_strength;
    energy_parts(p, 1.0, w, reinterpret_cast<const double*>(g1.data()),
                 [&](int s, int X, int Y) { return pd[(int64_t)s * d.n2 + (int64_t)X * d.N + Y]; }, corr, E);
    return 0;
}

int save_traj(Run& r, int i, const char* path) {
    if (!path || i < 0 || i >= r.ntraj) return G12_E_ARG;
    const DevState& d = r.ds[i];
    std::vector<double> g1(2 * d.n2), gxx(2 * d.n4), gxy(2 * d.n4);
    int rc = get_state(r, i, g1.data(), gxx.data(), gxy.data());
    if (rc) return rc;
    return g12io::save_checkpoint(path, d.Ns, d.Nb, r.t, r.steps, g1.data(), gxx.data(), gxy.data());
}

int load_traj(Run& r, int i, const char* path) {
    if (!path || i < 0 || i >= r.ntraj) return G12_E_ARG;
    const DevState& d = r.ds[i];
    std::vector<double> g1(2 * d.n2), gxx(2 * d.n4), gxy(2 * d.n4);
    double t = 0;
    int64_t steps = 0;
    int rc = g12io::load_checkpoint(path, d.Ns, d.Nb, &t, &steps, g1.data(), gxx.data(), gxy.data());
    if (rc) return rc;
    rc = set_state(r, i, g1.data(), gxx.data(), gxy.data());
    if (rc) return rc;
    r.t = t;
    r.steps = steps;
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
    // the G1 part of the stage set-up (h_HF of Y1 = S1 = G1x)
    RoutineBatch rb;
    if ((rc = rb.init(Ns, Nb, U, V, nV, false))) return rc;
    rb.tp.t1 = t1;
    rb.tp.t2 = t2;
    rb.tp.qstr = w;
    rb.tp.Nq = quench_extent;
    for (int i = 0; i < Nb; ++i) rb.tp.eps[i] = eps ? eps[i] : 0.0;
    if ((rc = rb.upload()) || (rc = up(rb.d.S1, G1x, sizeof(double2) * rb.d.n2))) return rc;
    StageRef sr = RoutineBatch::stage(1, 1.0, 0.0);
    sr.val.gate = 1;  // quench shift phi * strength with phi = 1: w on sites < quench_extent
    sr.val.phi = 1.0;
    g12::launch_b_stage_setup(rb.b, sr, G12_PREC_FP64, rb.s);
    std::vector<double> hf(2 * rb.d.n2);
    if ((rc = down(hf.data(), rb.d.h, sizeof(double2) * rb.d.n2))) return rc;
    for (int64_t e = 0; e < 2 * rb.d.n2; ++e) h[e] += hf[e];  // h accumulates
    return 0;
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
    if (!valid_precision_size(N, precision)) return G12_E_SIZE;
    const int64_t n2 = (int64_t)N * N, n4 = n2 * n2;
    Scratch sc;
    double2 *draw = sc.get<double2>(n4), *dT = sc.get<double2>(n4), *dC = sc.get<double2>(n4);
    double2 *dhx = sc.get<double2>(n2), *dhy = sc.get<double2>(n2);
    DEV_OR_FAIL(draw && dT && dC && dhx && dhy);
    if ((rc = up(draw, G2, sizeof(double2) * n4)) || (rc = up(dhx, h1x, sizeof(double2) * n2)) ||
        (rc = up(dhy, h1y, sizeof(double2) * n2)))
        return rc;
    g12::launch_ref_to_nat(draw, dT, Ns, Nb, api_stream());
    if ((rc = g12::launch_comm2(dT, dhx, dhy, dC, N, precision, api_stream())))
        return rc;
    g12::launch_nat_to_ref(dC, draw, Ns, Nb, api_stream());
    return down(comm, draw, sizeof(double2) * n4);
}

int g12_collision(double* Ix, const double* G2xyxy, const double* G2xxxx, int Ns, int Nb, double U,
                  const double* V, int nV) {
    if (!Ix || !G2xyxy || !G2xxxx || (nV > 0 && !V)) return G12_E_ARG;
    int rc = check_routine_geometry(Ns, Nb, nV);
    if (rc) return rc;
    // state spins (xxxx, xyxy), their partial traces, and the stage tables:
    // I[X][Y] = F^xx[X][Y][X]
    RoutineBatch rb;
    if ((rc = rb.init(Ns, Nb, U, V, nV, true)) || (rc = rb.upload())) return rc;
    if ((rc = rb.upload_nat(rb.d.S2[0], G2xxxx, rb.d.Y2[0])) || (rc = rb.upload_nat(rb.d.S2[1], G2xyxy, rb.d.Y2[0])))
        return rc;
    g12::launch_b_d_from_state(rb.b, rb.k, G12_PREC_FP64, rb.s);
    g12::launch_b_stage_setup(rb.b, RoutineBatch::stage(1, 1.0, 0.0), G12_PREC_FP64, rb.s);
    return down(Ix, rb.d.I, sizeof(double2) * rb.d.n2);
}

}  // extern "C"

// Psi of the stage G1 = G1 through the production epilogues: the strided pass
// writes its source terms into R (no products), the fused pass adds its own
// and takes the stage-0 update A = -i (R + sources); Psi = i A.  G2 = 0, so
// the partial traces and with them the F tables vanish and the sources are
// Psi's alone.  Spin 0 carries W_xx (Psi_xxxx), spin 1 W_xy (Psi_xyxy).
static int psi_common(double* psi, const double* G1, int Ns, int Nb, double U, const double* V, int nV,
                      bool same_spin) {
    if (!psi || !G1 || (nV > 0 && !V)) return G12_E_ARG;
    int rc = check_routine_geometry(Ns, Nb, nV);
    if (rc) return rc;
    RoutineBatch rb;
    if ((rc = rb.init(Ns, Nb, U, V, nV, true)) || (rc = rb.upload())) return rc;
    if ((rc = up(rb.d.S1, G1, sizeof(double2) * rb.d.n2))) return rc;
    const StageRef sr = RoutineBatch::stage(1, 1.0, 0.0);
    g12::launch_b_stage_setup(rb.b, sr, G12_PREC_FP64, rb.s);
    g12::launch_b_tile_args(rb.b, rb.k, sr, 0, 0, 1, 0, rb.s);
    g12::launch_b_tile_args(rb.b, rb.k, sr, 1, 0, 1, 0, rb.s);
    const int sp = same_spin ? 0 : 1;
    if ((rc = rb.download_ref(psi, rb.d.A2[sp], rb.d.R2[sp]))) return rc;
    times_i(psi, rb.d.n4);
    return 0;
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
    // Pi(P, Q) = g22 (F[p1][q1][p2] - F[p1][q1][q2]) with F built from
    // D^P + D^Q less the diagonal W term on D^Q: the production F table of the
    // spin whose state is Q.  Q as spin 0, P as spin 1, R = 0, the Qp and W
    // tables cleared after the set-up (so the sources are the F terms alone),
    // and the fused pass: A = -i Pi.
    RoutineBatch rb;
    if ((rc = rb.init(Ns, Nb, U, V, nV, true)) || (rc = rb.upload())) return rc;
    if ((rc = up(rb.d.S1, G1y, sizeof(double2) * rb.d.n2)) ||
        (rc = rb.upload_nat(rb.d.S2[0], G2xxxx, rb.d.Y2[0])) || (rc = rb.upload_nat(rb.d.S2[1], G2xyxy, rb.d.Y2[0])))
        return rc;
    const StageRef sr = RoutineBatch::stage(1, 1.0, 0.0);
    g12::launch_b_d_from_state(rb.b, rb.k, G12_PREC_FP64, rb.s);
    g12::launch_b_stage_setup(rb.b, sr, G12_PREC_FP64, rb.s);
    DevState& d = rb.d;
    for (void* z : {(void*)d.Qp[0], (void*)d.Qp[1]})
        if (hipMemsetAsync(z, 0, sizeof(double2) * d.n3, rb.s) != hipSuccess) return G12_E_HIP;
    for (void* z : {(void*)d.Wxy, (void*)d.Wxx})
        if (hipMemsetAsync(z, 0, sizeof(double) * d.n2, rb.s) != hipSuccess) return G12_E_HIP;
    g12::launch_b_tile_args(rb.b, rb.k, sr, 1, 0, 1, 0, rb.s);
    if ((rc = rb.download_ref(pi, rb.d.A2[0], rb.d.R2[0]))) return rc;
    times_i(pi, rb.d.n4);
    return 0;
}

int g12_h2_commutator_noalloc(const double* G2, const double* h1x, const double* h1y, double* comm, int Ns, int Nb,
                              int precision, void* stream) {
    if (!G2 || !h1x || !h1y || !comm || Ns < 1 || Nb < 1 || Nb > G12_MAX_NB) return G12_E_ARG;
    if (precision < G12_PREC_FP64 || precision > G12_PREC_BF16X3) return G12_E_ARG;
    const int N = Ns * Nb;
    if (N > G12_NOALLOC_MAX_N) {
        std::fprintf(stderr,
                     "g12_h2_commutator_noalloc: N = Ns*Nb = %d exceeds the no-allocation limit %d "
                     "(use g12_h2_commutator)\n",
                     N, G12_NOALLOC_MAX_N);
        return G12_E_SIZE;
    }
    // the thread's last-error state may hold an earlier, unrelated call's
    // result (a stream or event query that returned "not ready"); the launch
    // checks below must see only this call's launches
    (void)hipGetLastError();
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        std::fprintf(stderr, "g12_h2_commutator_noalloc: hipGetDevice failed: %s\n", hipGetErrorString(e));
        return G12_E_HIP;
    }
    char* base = static_cast<char*>(scratch_address(dev));
    if (!base) return G12_E_HIP;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t n4 = (int64_t)N * N * N * N;
    // scratch: descriptors, natural-layout input, natural-layout result
    double2* nat_in = reinterpret_cast<double2*>(base + g12::COMM2_DESC_BYTES);
    double2* nat_out = nat_in + n4;
    g12::launch_ref_to_nat(reinterpret_cast<const double2*>(G2), nat_in, Ns, Nb, s);
    (void)g12::enqueue_comm2(nat_in, reinterpret_cast<const double2*>(h1x), reinterpret_cast<const double2*>(h1y),
                             nat_out, N, precision, base, s);
    g12::launch_nat_to_ref(nat_out, reinterpret_cast<double2*>(comm), Ns, Nb, s);
    e = hipGetLastError();
    if (e != hipSuccess) {
        std::fprintf(stderr, "g12_h2_commutator_noalloc: launch failed: %s\n", hipGetErrorString(e));
        return G12_E_HIP;
    }
    return 0;
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
    // R = comm + psi + pi - pi^dagger-partner, then the fused epilogue with no
    // products or sources at a middle stage of weight w: A = RK + w K and the
    // next stage input S + 1 K = K (S = 0), K = -i R.  Spin 0 carries the
    // data; spin 1's buffers serve as staging.
    RoutineBatch rb;
    if ((rc = rb.init(Ns, Nb, 0.0, nullptr, 0, true)) || (rc = rb.upload())) return rc;
    DevState& d = rb.d;
    double2* raw = d.S2[1];
    if ((rc = rb.upload_nat(d.A2[1], comm, raw)) || (rc = rb.upload_nat(d.R2[1], psi, raw)) ||
        (rc = rb.upload_nat(d.Y2[1], pi, raw)) || (rc = rb.upload_nat(d.A2[0], RK_result, raw)))
        return rc;
    g12::launch_g2_sum(d.A2[1], d.R2[1], d.Y2[1], d.R2[0], d.N, rb.s);
    g12::launch_b_tile_args(rb.b, rb.k, RoutineBatch::stage(0, w, 1.0), 1, 0, 0, 0, rb.s);
    if ((rc = rb.download_ref(K, d.Y2[0], raw))) return rc;
    return rb.download_ref(RK_result, d.A2[0], raw);
}

// ---------------------------------------------------------------------------
// energy (see energy_parts)

int g12_energy(const g12_params* p, double lam, double w, const double* G1, const double* G2xx, const double* G2xy,
               double* E) {
    if (!p || !G1 || !E || !valid_geometry(p->Ns, p->Nb)) return G12_E_ARG;
    typedef std::complex<double> cplx;
    const int Ns = p->Ns, Nb = p->Nb, N = Ns * Nb;
    const int64_t N2 = (int64_t)N * N;
    const cplx* M[2] = {reinterpret_cast<const cplx*>(G2xx), reinterpret_cast<const cplx*>(G2xy)};
    const bool corr = G2xx && G2xy;
    // reference pair index of (X, Y) = ((i, a), (j, b)): i + j Ns + (a + b Nb) Ns^2
    auto pair = [&](int X, int Y) {
        return (int64_t)(X % Ns) + (int64_t)(Y % Ns) * Ns + ((int64_t)(X / Ns) + (int64_t)(Y / Ns) * Nb) * Ns * Ns;
    };
    energy_parts(*p, lam, w, G1,
                 [&](int s, int X, int Y) {
                     const int64_t A = pair(X, Y);
                     return M[s][A * N2 + A];
                 },
                 corr, E);
    return 0;
}

int g12_state_energy(g12_ctx* c, double* E) { return c ? state_energy(c->r, 0, E) : G12_E_ARG; }

int g12_batch_state_energy(g12_batch* b, int i, double* E) { return b ? state_energy(b->r, i, E) : G12_E_ARG; }

// ---------------------------------------------------------------------------
// checkpoints: magic, N, Ns, Nb, t, steps, G1, G2xx, G2xy (reference layouts)

int g12_save_state(g12_ctx* c, const char* path) { return c ? save_traj(c->r, 0, path) : G12_E_ARG; }

int g12_load_state(g12_ctx* c, const char* path) { return c ? load_traj(c->r, 0, path) : G12_E_ARG; }

int g12_batch_save_state(g12_batch* b, int i, const char* path) { return b ? save_traj(b->r, i, path) : G12_E_ARG; }

int g12_batch_load_state(g12_batch* b, int i, const char* path) { return b ? load_traj(b->r, i, path) : G12_E_ARG; }

}  // extern "C"
