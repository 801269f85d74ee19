// HIP kernels for the HF-GKBA G1-G2 (GW) right-hand side on gfx950.
//
// Notation (all complex numbers are double2 = (re, im)):
//   N = Ns*Nb, one-particle index p = site + orb*Ns.
//   G2 tensors are held in the natural layout T[p1][p2][q1][q2]
//   (= M[(p1,p2);(q1,q2)] of the reference, pair order permuted once at
//   upload; see launch_ref_to_nat).
//
// The two-particle commutator [H2, M] (H2 = h(x)1 + 1(x)h, reference
// commutators.h:53-99) is evaluated as two passes over N x N tiles, each
// computing the one-particle commutator  O = h.X - X.h  of a tile X:
//   strided tiles   X[p1][q1] = T[p1][p2][q1][q2] at fixed (p2,q2)
//                   -> the terms that act on p1 and q1;
//   contiguous tiles X[p2][q2] = T[p1][p2][q1][q2] at fixed (p1,q1)
//                   -> the terms that act on p2 and q2.
// Between the two passes a streaming kernel adds the rest of the G2
// right-hand side (Psi, Pi, Pi-dagger, reference G1_G2_t_step.h:55-86) into
// the strided-pass result, so Psi and Pi are never stored on their own; the
// contiguous pass then finishes K and applies the RK4 update for its tile.
#include "g12_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

namespace g12 {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cdiv6(double2 a) { return make_double2(a.x / 6.0, a.y / 6.0); }
// i*z and -i*z
__device__ __forceinline__ double2 cmul_i(double2 a) { return make_double2(-a.y, a.x); }
__device__ __forceinline__ double2 cmul_mi(double2 a) { return make_double2(a.y, -a.x); }

__device__ __forceinline__ bool failed(const int* fail) { return fail && *fail != 0; }

// Explicit global-address-space accesses: pointers read from the trajectory
// descriptors are generic, and generic (flat) memory operations are also
// counted by the LDS wait counter, so in the tile kernels every LDS wait of the
// MFMA loop would wait for the next tile's prefetch and every LDS barrier for
// the previous tile's stores.
#define GLOBAL __attribute__((address_space(1)))
typedef double dv2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double2 gload(const double2* p, int64_t i) {
    const dv2 v = ((const GLOBAL dv2*)p)[i];
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void gstore(double2* p, int64_t i, double2 v) {
    ((GLOBAL dv2*)p)[i] = dv2{v.x, v.y};
}
__device__ __forceinline__ double gload1(const double* p, int64_t i) { return ((const GLOBAL double*)p)[i]; }

// Complex elements stored as complex128 or, in the compact (complex64) form
// of mixed-precision runs, as complex64; arithmetic stays FP64.
typedef float fv2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 cload(const void* p, int64_t i, bool compact) {
    if (compact) {
        const fv2 v = ((const GLOBAL fv2*)p)[i];
        return make_double2(v.x, v.y);
    }
    return gload((const double2*)p, i);
}
__device__ __forceinline__ void cstore(void* p, int64_t i, double2 v, bool compact) {
    if (compact)
        ((GLOBAL fv2*)p)[i] = fv2{(float)v.x, (float)v.y};
    else
        gstore((double2*)p, i, v);
}

// The same accesses with 32-bit element offsets from a base pointer held in
// scalar registers.  The compiler emits them in the global saddr form (scalar
// base + 32-bit byte offset per lane), one VGPR per address instead of two
// and no 64-bit address arithmetic.  Valid for buffers below 4 GB: the tile
// kernels index N^4 complex128 elements with N <= 64 (268 MB).
__device__ __forceinline__ double2 gld(const double2* p, uint32_t i) {
    const dv2 v = *(const GLOBAL dv2*)((const GLOBAL char*)p + (i << 4));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void gst(double2* p, uint32_t i, double2 v) {
    *(GLOBAL dv2*)((GLOBAL char*)p + (i << 4)) = dv2{v.x, v.y};
}
__device__ __forceinline__ double gld1(const double* p, uint32_t i) {
    return *(const GLOBAL double*)((const GLOBAL char*)p + (i << 3));
}
__device__ __forceinline__ double2 cld(const void* p, uint32_t i, bool compact) {
    if (compact) {
        const fv2 v = *(const GLOBAL fv2*)((const GLOBAL char*)p + (i << 3));
        return make_double2(v.x, v.y);
    }
    return gld((const double2*)p, i);
}
__device__ __forceinline__ void cst(void* p, uint32_t i, double2 v, bool compact) {
    if (compact)
        *(GLOBAL fv2*)((GLOBAL char*)p + (i << 3)) = fv2{(float)v.x, (float)v.y};
    else
        gst((double2*)p, i, v);
}

// Non-temporal forms for data streamed once per pass (the pair-symmetric
// product pass's result R, the update pass's X and outputs): measured 4-6%
// more steps/s at N = 64 than the same accesses through the normal cache
// policy; non-temporal state loads measured no faster.  Complex64 buffers use
// them only where rows fill whole 128-byte lines (N a multiple of 16): with
// the ensemble's N = 6..60 they made the FP32-3M batch 5% slower.
__host__ __device__ inline bool use_nt(bool compact, int N) { return !compact || (N & 15) == 0; }
__device__ __forceinline__ void cst_nt(void* p, uint32_t i, double2 v, bool compact) {
    if (compact)
        __builtin_nontemporal_store(fv2{(float)v.x, (float)v.y}, (GLOBAL fv2*)((GLOBAL char*)p + (i << 3)));
    else
        __builtin_nontemporal_store(dv2{v.x, v.y}, (GLOBAL dv2*)((GLOBAL char*)p + (i << 4)));
}

// Values that are the same in every lane, moved to scalar registers: the
// trajectory descriptors are read with vector loads (the kernel writes global
// memory, so scalar loads of them are not allowed), and left in vector
// registers they would crowd out the tile data.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t uni64(int64_t x) {
    const uint64_t u = (uint64_t)x;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <class T>
__device__ __forceinline__ T* uni_ptr(T* p) {
    return reinterpret_cast<T*>(uni64((int64_t)reinterpret_cast<uintptr_t>(p)));
}
__device__ __forceinline__ double uni_d(double x) {
    return __builtin_bit_cast(double, uni64(__builtin_bit_cast(int64_t, x)));
}


// Workgroup barrier that orders LDS only: the tile loop reuses its LDS buffers
// but never exchanges global memory between waves, so waiting for the
// scattered global stores of the previous tile (as a full __syncthreads fence
// would) only stalls the MFMA pipeline.
__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------------------
// Interaction tables W_xy / W_xx (N x N):  U on one site, V_n at distance
// n <= nV, 0 beyond; W_xx(X,X) = 0.  (The loops of HF.h, psi.h, pi.h and
// collision_integral.h over r and n enumerate exactly these entries.)
__device__ inline double site_w(int sx, int sz, const StageParams& sp) {  // W_xy of sites sx, sz
    int d = sx - sz;
    d = d < 0 ? -d : d;
    if (d == 0) return sp.U;
    return d <= sp.nV ? sp.V[d - 1] : 0.0;
}
__device__ inline double w_xy(int x, int z, int Ns, const StageParams& sp) { return site_w(x % Ns, z % Ns, sp); }

// ---------------------------------------------------------------------------
// h = h_HF[G] for one element (X, Y) (h_HFx_quench, HF.h:57-104):
//   diagonal: eps + quench + Hartree  -iU g_XX - 2i sum_{Z != X} W_xy(X,Z) g_ZZ
//   -t1 between neighbouring sites of one orbital, -t2 between adjacent
//   orbitals of one site, onsite Fock +iU g_YX (transposed), long-range Fock
//   +iV g_XY.
// Formed during the stage set-up: the Hartree sums read the diagonal of the
// stage G1 from LDS, the Fock terms take G1 elements from gv(index).
template <class GV>
__device__ inline double2 hf_element_with(GV gv, const double2* diag, int X, int Y, int Ns, int Nb,
                                          const StageParams& sp) {
    const int N = Ns * Nb;
    const int sx = X % Ns, ox = X / Ns, sy = Y % Ns, oy = Y / Ns;
    double2 v = make_double2(0.0, 0.0);
    if (X == Y) {
        v.x += sp.eps[ox];
        if (sx < sp.Nq) v.x += sp.wq;
        double2 acc = cscale(diag[X], sp.U);  // -i U g_XX
        double2 lr = make_double2(0.0, 0.0);
        for (int oz = 0, Z = 0; oz < Nb; ++oz)
            for (int sz = 0; sz < Ns; ++sz, ++Z) {
                if (Z == X) continue;
                const double w = site_w(sx, sz, sp);
                if (w != 0.0) lr = cadd(lr, cscale(diag[Z], w));
            }
        acc = cadd(acc, cscale(lr, 2.0));
        v = cadd(v, cmul_mi(acc));
    } else {
        int ds = sx - sy;
        ds = ds < 0 ? -ds : ds;
        if (ox == oy && ds == 1) v.x -= sp.t1;
        if (sx == sy) {
            int dorb = ox - oy;
            dorb = dorb < 0 ? -dorb : dorb;
            if (dorb == 1) v.x -= sp.t2;
            v = cadd(v, cmul_i(cscale(gv((int64_t)Y * N + X), sp.U)));
        } else if (ds <= sp.nV) {
            v = cadd(v, cmul_i(cscale(gv((int64_t)X * N + Y), sp.V[ds - 1])));
        }
    }
    return v;
}

// Per-stage scalars of trajectory t from its constants and the shared stage
// block (the reference's stage set-up, RK_solver.h:120-151 / AS.h:111-141).
__device__ inline StageShared stage_of(const StageRef& r) {
    if (r.base) return r.base[(*r.counter) * 4 + r.stage];
    return r.val;
}

// Stage parameters of trajectory t, filled by the threads of a block
// together (thread i writes shell i).  Callers synchronise before use.
__device__ inline void fill_sp(StageParams& sp, const TrajParams& t, const StageShared& ss) {
    const double U = ss.as_phase ? ss.lam * t.U : t.U;  // U_AS = fermi(t) U during switching
    for (int i = threadIdx.x; i < t.nV; i += blockDim.x) sp.V[i] = t.v_abs ? t.Vprof[i] : U * t.Vprof[i];
    if (threadIdx.x == 0) {
        sp.c = ss.c;
        sp.wk = ss.wk;
        sp.dt = ss.dt;
        sp.first = ss.first;
        sp.last = ss.last;
        sp.corr = t.corr;
        sp.U = U;
        sp.nV = t.nV;
        sp.wq = ss.gate ? ss.phi * t.qstr : 0.0;
        sp.Nq = ss.gate ? t.Nq : 0;
        sp.t1 = t.t1;
        sp.t2 = t.t2;
    }
    if (threadIdx.x < G12_MAX_NB) sp.eps[threadIdx.x] = t.eps[threadIdx.x];
}

// A trajectory takes part in a stage unless its guard tripped or its own
// adiabatic switching is already complete.
__device__ inline bool active(const DevState& d, const TrajParams& t, const StageShared& ss) {
    if (failed(d.fail)) return false;
    return !(ss.as_phase && ss.step >= t.as_steps);
}

// Stage set-up for G1: Y1 = S1 + c K1 and its conjugate transpose G1t (read
// by the strided pass), h = h_HF[Y1], W tables.  Blocks of blockDim elements
// (`chunk`) per trajectory; each block forms the diagonal of Y1 itself (for
// the Hartree sums, in `diag`) and takes the Y1 elements of its Fock terms
// from S1 and K1, so the blocks need no ordering among themselves.
__device__ inline void stage_g1_block(const DevState& d, const StageShared& ss, const StageParams& sp, int chunk,
                                      double2* diag) {
    const int N = d.N;
    const int n2 = N * N;
    const int e0 = chunk * blockDim.x;
    const double2* S1 = uni_ptr(d.S1);
    const double2* K1 = uni_ptr(d.K1);
    const bool first = ss.first != 0;
    const double c = ss.c;
    auto yv = [&](int64_t i) {
        double2 y = gload(S1, i);
        if (!first) y = cadd(y, cscale(gload(K1, i), c));
        return y;
    };
    for (int X = threadIdx.x; X < N; X += blockDim.x) diag[X] = yv((int64_t)X * N + X);
    __syncthreads();
    const int e = e0 + threadIdx.x;
    if (e < n2) {
        const int X = e / N, Y = e % N;
        const double2 y = yv(e);
        gstore(d.Y1, e, y);
        gstore(d.G1t, e, cconj(yv((int64_t)Y * N + X)));
        gstore(d.h, e, hf_element_with(yv, diag, X, Y, d.Ns, d.Nb, sp));
        const double w = w_xy(X, Y, d.Ns, sp);
        ((GLOBAL double*)d.Wxy)[e] = w;
        ((GLOBAL double*)d.Wxx)[e] = (X == Y) ? 0.0 : w;
    }
}

// G1 right-hand side: comm = [h, Y1], K = -i(comm + I + I^dag), RK
// accumulation (G1_step, G1_G2_t_step.h:24-47) and the final update.  One
// block per row X: the rows X of h and Y1 are staged in LDS, and four groups
// of 64 threads each sum a quarter of the z range for the columns Y they own.
__global__ __launch_bounds__(256) void k_g1_rhs(BatchDev b, StageRef r) {
    const int ti = blockIdx.y;
    const DevState& d = b.ds[ti];
    const TrajParams& tp = b.tp[ti];
    const StageShared ss = stage_of(r);
    if (!active(d, tp, ss)) return;
    const int N = d.N;
    const int X = blockIdx.x;
    if (X >= N) return;
    __shared__ double2 hrow[G12_MAX_N], grow[G12_MAX_N];
    __shared__ double2 part[4][64];
    const double2* h = d.h;
    const double2* Y1 = d.Y1;
    for (int z = threadIdx.x; z < N; z += blockDim.x) {
        hrow[z] = gload(h, (int64_t)X * N + z);
        grow[z] = gload(Y1, (int64_t)X * N + z);
    }
    __syncthreads();
    const int grp = threadIdx.x >> 6, col = threadIdx.x & 63;
    const int zq = (N + 3) / 4, z0 = grp * zq, z1 = min(N, z0 + zq);
    for (int Y0 = 0; Y0 < N; Y0 += 64) {
        const int Y = Y0 + col;
        double2 acc = make_double2(0.0, 0.0);
        if (Y < N) {
#pragma unroll 4
            for (int z = z0; z < z1; ++z) {
                acc = cadd(acc, cmul(hrow[z], gload(Y1, (int64_t)z * N + Y)));
                acc = csub(acc, cmul(grow[z], gload(h, (int64_t)z * N + Y)));
            }
        }
        part[grp][col] = acc;
        __syncthreads();
        if (grp == 0 && Y < N) {
            const int64_t e = (int64_t)X * N + Y;
            double2 v = cadd(cadd(part[0][col], part[1][col]), cadd(part[2][col], part[3][col]));
            if (tp.corr) {
                const double2 iyx = cconj(d.I[(int64_t)Y * N + X]);
                // the collision term J = I + I^dagger of a production step's
                // first stage (whose input is the state) goes to the record ring
                if (ss.first && !ss.as_phase && d.jtraj) d.jtraj[(int64_t)ss.slot * N * N + e] = cadd(d.I[e], iyx);
                v = cadd(cadd(v, d.I[e]), iyx);
            } else if (d.jext && *d.jext_on) {
                // a given collision term at the step's two ends: stage 0 at t,
                // stages 1 and 2 at t + dt/2 (their mean), stage 3 at t + dt
                const double2* J = d.jext + (int64_t)ss.slot * N * N;
                const double2 a = J[e], b = J[(int64_t)N * N + e];
                v = cadd(v, ss.first ? a : ss.last ? b : cscale(cadd(a, b), 0.5));
            }
            const double2 K = cmul_mi(v);
            const double2 wK = cscale(K, ss.wk);
            if (ss.first) {
                d.A1[e] = wK;
                d.K1[e] = K;
            } else if (!ss.last) {
                d.A1[e] = cadd(d.A1[e], wK);
                d.K1[e] = K;
            } else {
                const double2 A = cadd(d.A1[e], wK);
                d.S1[e] = cadd(d.S1[e], cdiv6(cscale(A, ss.dt)));
            }
        }
        __syncthreads();
    }
}

// Records G1 before a production step and applies the density guard
// (RK_solver.h:100-110): |2 Im G1[0,0]| > 3 or NaN stops the trajectory after
// the record; fail = step + 1.  Step index and ring slot come from the stage-0
// block of the step.
__global__ void k_record(BatchDev b, StageRef r) {
    const int ti = blockIdx.y;
    const DevState& d = b.ds[ti];
    if (failed(d.fail)) return;
    const StageShared ss = stage_of(r);
    const int n2 = d.N * d.N;
    // one element per thread over the blocks of the trajectory's row of the grid
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (d.traj && e < n2) gstore(d.traj + (int64_t)ss.slot * n2, e, gload(d.S1, e));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const double dens = 2 * d.S1[0].y;
        if (fabs(dens) > 3 || isnan(dens)) *d.fail = ss.step + 1;
    }
}

__global__ void k_advance_counter(int* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter += 1;
}

// ---------------------------------------------------------------------------
// Tile commutator on FP64 MFMA.
//
// Each tile is an N x N slice X of a G2 tensor (padded to NP, a multiple of
// 16); the kernel computes O = h.X - X.h with v_mfma_f64_16x16x4f64, three real
// MFMAs per complex product (3M), and 16x16 output blocks dealt round-robin to
// the waves.  Workgroups are persistent: h stays in LDS for the whole launch, and
// while a tile is multiplied the next tile's elements are already loading into
// registers.  MFMA operand layout (16x16x4 f64): A[m][k] in lane m + 16k,
// B[k][n] in lane n + 16k; result register j of lane l holds
// D[4j + l/16][l%16].
template <int NP>
struct TileLds {
    static constexpr int P = NP + 1;  // padded pitch (doubles)
    double xr[NP * P];
    double xi[NP * P];
    double hr[NP * P];
    double hi[NP * P];
};

template <int NP, int NW_>
struct TileCfg {
    static constexpr int NWAVE = NW_;
    static constexpr int NTHR = NWAVE * 64;
    static constexpr int NT = NP / 16;
    static constexpr int TT = NT * NT;
    static constexpr int PW = (TT + NWAVE - 1) / NWAVE;  // output blocks per wave
    static constexpr int PF = (NP * NP + NTHR - 1) / NTHR;  // prefetched elements per thread
};

template <int NP, int NW>
__device__ __forceinline__ void load_frags(const double* __restrict__ ar_, const double* __restrict__ ai_,
                                           const double* __restrict__ br_, const double* __restrict__ bi_, int kk,
                                           int wave, int r, double (&a_r)[TileCfg<NP, NW>::PW],
                                           double (&a_i)[TileCfg<NP, NW>::PW], double (&b_r)[TileCfg<NP, NW>::PW],
                                           double (&b_i)[TileCfg<NP, NW>::PW]) {
    constexpr int P = TileLds<NP>::P;
    constexpr int NT = TileCfg<NP, NW>::NT, TT = TileCfg<NP, NW>::TT, PW = TileCfg<NP, NW>::PW;
    constexpr bool FULL = (TT % NW) == 0;  // every wave owns PW blocks
#pragma unroll
    for (int u = 0; u < PW; ++u) {
        const int t = wave + NW * u;
        const bool on = FULL || t < TT;
        const int mm = (t / NT) * 16 + r, nn = (t % NT) * 16 + r;
        a_r[u] = on ? ar_[mm * P + kk] : 0.0;
        a_i[u] = on ? ai_[mm * P + kk] : 0.0;
        b_r[u] = on ? br_[kk * P + nn] : 0.0;
        b_i[u] = on ? bi_[kk * P + nn] : 0.0;
    }
}

// Complex tile product on three real MFMAs per k-step (3M): with As = Ar + Ai
// and Bs = Br + Bi formed in registers,
//   P1 += Ar Br, P2 += Ai Bi, P3 += As Bs;  Re = P1 - P2, Im = P3 - P1 - P2
// (sign NEG: -A.B).  The rounding error of Im is that of a sum of
// |A||B|-sized terms, the order of the cancellation between h.X and X.h that
// any product scheme carries; against four real MFMAs per k-step the N = 64
// workload's densities differ by 2e-14 after 5305 steps, for 25% fewer MFMAs.
// Lane group kq (= lane / 16) of the 16x16x4 MFMA takes k = kq*NP/4 + j at
// step j: with the NP+1 pitch this puts the two lane groups that share an LDS
// cycle 16 rows / columns apart, so A and B fragment reads are conflict-free
// at NP = 64 (the sum over k does not depend on its order).
template <int NP, int NW, bool NEG>
__device__ __forceinline__ void tile_product_3m(const double* ar_, const double* ai_, const double* br_,
                                                const double* bi_, d4 (&p1)[TileCfg<NP, NW>::PW],
                                                d4 (&p2)[TileCfg<NP, NW>::PW], d4 (&p3)[TileCfg<NP, NW>::PW],
                                                int wave, int lane) {
    constexpr int PW = TileCfg<NP, NW>::PW, KS = NP / 4;
    const int r = lane & 15, kb = (lane >> 4) * KS;
#pragma unroll 4
    for (int j = 0; j < KS; ++j) {
        double a_r[PW], a_i[PW], b_r[PW], b_i[PW];
        load_frags<NP, NW>(ar_, ai_, br_, bi_, kb + j, wave, r, a_r, a_i, b_r, b_i);
#pragma unroll
        for (int u = 0; u < PW; ++u)
            p1[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(NEG ? -a_r[u] : a_r[u], b_r[u], p1[u], 0, 0, 0);
#pragma unroll
        for (int u = 0; u < PW; ++u)
            p2[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(NEG ? -a_i[u] : a_i[u], b_i[u], p2[u], 0, 0, 0);
#pragma unroll
        for (int u = 0; u < PW; ++u) {
            const double as = a_r[u] + a_i[u], bs = b_r[u] + b_i[u];
            p3[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(NEG ? -as : as, bs, p3[u], 0, 0, 0);
        }
    }
}

template <int NP, int NW, class Mid>
__device__ __forceinline__ void tile_mfma_fp64(const TileLds<NP>& L, int hmode, d4 (&ar)[TileCfg<NP, NW>::PW],
                                               d4 (&ai)[TileCfg<NP, NW>::PW], int wave, int lane, Mid&& mid) {
    constexpr int PW = TileCfg<NP, NW>::PW;
    d4 p1[PW], p2[PW], p3[PW];
#pragma unroll
    for (int u = 0; u < PW; ++u) {
        p1[u] = d4{0.0, 0.0, 0.0, 0.0};
        p2[u] = p1[u];
        p3[u] = p1[u];
    }
    if (hmode & 1) tile_product_3m<NP, NW, false>(L.hr, L.hi, L.xr, L.xi, p1, p2, p3, wave, lane);  // + h.X
    // v enters Re = P1 - P2 through P1 and Im = P3 - P1 - P2 through P3
    mid([&](int u, int j, double2 v) {
        p1[u][j] += v.x;
        p3[u][j] += v.x + v.y;
    });
    if (hmode & 2) tile_product_3m<NP, NW, true>(L.xr, L.xi, L.hr, L.hi, p1, p2, p3, wave, lane);   // - X.h
#pragma unroll
    for (int u = 0; u < PW; ++u) {
        ar[u] = p1[u] - p2[u];
        ai[u] = p3[u] - p1[u] - p2[u];
    }
}

// ---------------------------------------------------------------------------
// Stage tables on FP64 MFMA.  One block per (trajectory, X, 64 x 64 output
// block) and table family forms its block of an N x N x N complex product with
// the tile product of the commutator passes, the sum over w taken in chunks of
// 64 (operands zero-padded; one chunk and one output block per X at N <= 64):
//   Qp: O[Z][Y] = sum_w A[Z][w] Y1[w][Y],  A[Z][w] = Y1[X][w] W_xy[Z][w];
//       Qp^xy[X][Y][Z] = O[Z][Y],  Qp^xx[X][Y][Z] = O[Z][Y] - A[Z][Z] Y1[Z][Y]
//       (W_xx equals W_xy off the diagonal and is 0 on it);
//   F:  S[Y][Z] = sum_w (D^xx + D^xy)[X][Y][w] W_xy[w][Z],
//       F^xx[X][Y][Z] = -i (S[Y][Z] - W_xy[Z][Z] D^xx[X][Y][Z]),
//       F^xy[X][Y][Z] = -i (S[Y][Z] - W_xy[Z][Z] D^xy[X][Y][Z]),
//       I[X][Y] = F^xx[X][Y][X]
// with Qp^s[X][Y][Z] = Q^s[Z][X][Y] = sum_w Y1[X][w] W^s[Z][w] Y1[w][Y]
// (psi.h:26-277), F as in Pi_xyxy with the swapped G2 pair (pi.h:25-95,
// RK_solver.h:163-164) and I as in collision_int (collision_integral.h:27-69).
// The stage G1 Y1 = S1 + c K1 and W are formed here as the G1 part forms
// them.  With `herm` (mixed-precision runs) D enters through its Hermitian
// part (D[X][Y][Z] + conj D[Y][X][Z]) / 2.  A Hermitian G2 (the physical
// states: G2 evolved from zero stays Hermitian) has D[X][Y][Z] =
// conj D[Y][X][Z], and that identity makes Re tr I, and with it the change of
// tr G1, vanish; FP32 or BF16 tile products break it slightly, the average
// restores particle-number conservation to FP64 level.  FP64 runs use D as it
// is, as the reference does, for any input.
constexpr int TG = G12_TABLE_GROUP;
static_assert(TG == 64, "table output blocks of 64 x 64");
constexpr int SETUP_THREADS = 1024;
typedef TileCfg<64, 16> SetupCfg;
static_assert(SetupCfg::NTHR == SETUP_THREADS && SetupCfg::PW == 1, "one 16x16 output block per wave");

// (re, im) of this wave's 16x16 block of sum over chunks of A.B: fill(chunk)
// writes A into (hr, hi) and B into (xr, xi); register j of lane l holds row
// 4j + l/16, column l%16 of the block
template <class Fill>
__device__ inline void setup_product(TileLds<64>& L, int nchunks, Fill&& fill, d4& re, d4& im, int wave,
                                     int lane) {
    d4 p1[1] = {d4{0.0, 0.0, 0.0, 0.0}}, p2[1] = {p1[0]}, p3[1] = {p1[0]};
    for (int c = 0; c < nchunks; ++c) {
        if (c) __syncthreads();  // the previous chunk's readers are done
        fill(c);
        __syncthreads();
        tile_product_3m<64, 16, false>(L.hr, L.hi, L.xr, L.xi, p1, p2, p3, wave, lane);
    }
    re = p1[0] - p2[0];
    im = p3[0] - p1[0] - p2[0];
}

// Qp block (rows Z in block zb, columns Y in block yb) of row X
__device__ inline void q_tables_block(const DevState& d, const StageShared& ss, const StageParams& sp, int X,
                                      int zb, int yb, TileLds<64>& L) {
    constexpr int P = TileLds<64>::P;
    const int N = d.N;
    const double2* S1 = uni_ptr(d.S1);
    const double2* K1 = uni_ptr(d.K1);
    auto y1 = [&](int64_t i) {  // stage G1
        double2 y = gload(S1, i);
        if (!ss.first) y = cadd(y, cscale(gload(K1, i), ss.c));
        return y;
    };
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    d4 re, im;
    setup_product(L, (N + TG - 1) / TG, [&](int c) {
        for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {  // B = Y1 rows w, columns Y
            const int w = c * TG + (e >> 6), Y = yb * TG + (e & 63);
            const double2 y = (w < N && Y < N) ? y1((int64_t)w * N + Y) : make_double2(0.0, 0.0);
            L.xr[(e >> 6) * P + (e & 63)] = y.x;
            L.xi[(e >> 6) * P + (e & 63)] = y.y;
        }
        for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {  // A[Z][w] = Y1[X][w] W_xy[Z][w]
            const int Z = zb * TG + (e >> 6), w = c * TG + (e & 63);
            double2 a = make_double2(0.0, 0.0);
            if (Z < N && w < N) a = cscale(y1((int64_t)X * N + w), w_xy(w, Z, d.Ns, sp));
            L.hr[(e >> 6) * P + (e & 63)] = a.x;
            L.hi[(e >> 6) * P + (e & 63)] = a.y;
        }
    }, re, im, wave, lane);
    const int Y = yb * TG + (wave & 3) * 16 + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int Z = zb * TG + (wave >> 2) * 16 + 4 * j + (lane >> 4);
        if (Z >= N || Y >= N) continue;
        const double2 o = make_double2(re[j], im[j]);
        const double2 wzz = cmul(cscale(y1((int64_t)X * N + Z), w_xy(Z, Z, d.Ns, sp)),
                                 y1((int64_t)Z * N + Y));  // the w = Z term
        const int64_t qp = ((int64_t)X * N + Y) * N + Z;
        d.Qp[0][qp] = csub(o, wzz);
        d.Qp[1][qp] = o;
    }
}

// F block (rows Y in block yb, columns Z in block zb) of row X
__device__ inline void f_tables_block(const DevState& d, const StageParams& sp, int X, int yb, int zb, bool herm,
                                      TileLds<64>& L) {
    constexpr int P = TileLds<64>::P;
    const int N = d.N;
    auto dval = [&](int s, int Y, int w) {  // D^s[X][Y][w], its Hermitian part with herm
        double2 v = d.D[s][((int64_t)X * N + Y) * N + w];
        if (herm) v = cscale(cadd(v, cconj(d.D[s][((int64_t)Y * N + X) * N + w])), 0.5);
        return v;
    };
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    d4 re, im;
    setup_product(L, (N + TG - 1) / TG, [&](int c) {
        for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {  // A = D^xx + D^xy, B = W_xy (real)
            const int m = e >> 6, n = e & 63;
            const int Y = yb * TG + m, w = c * TG + n;     // A[Y][w]
            const int wb = c * TG + m, Z = zb * TG + n;    // B[w][Z]
            double2 a = make_double2(0.0, 0.0);
            if (Y < N && w < N) a = cadd(dval(0, Y, w), dval(1, Y, w));
            L.hr[m * P + n] = a.x;
            L.hi[m * P + n] = a.y;
            L.xr[m * P + n] = (wb < N && Z < N) ? w_xy(wb, Z, d.Ns, sp) : 0.0;
            L.xi[m * P + n] = 0.0;
        }
    }, re, im, wave, lane);
    const int Z = zb * TG + (wave & 3) * 16 + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int Y = yb * TG + (wave >> 2) * 16 + 4 * j + (lane >> 4);
        if (Y >= N || Z >= N) continue;
        const double2 sv = make_double2(re[j], im[j]);
        const double wzz = w_xy(Z, Z, d.Ns, sp);  // the w = Z weight W_xx lacks
        const double2 fxx = cmul_mi(csub(sv, cscale(dval(0, Y, Z), wzz)));
        const double2 fxy = cmul_mi(csub(sv, cscale(dval(1, Y, Z), wzz)));
        const int64_t f = ((int64_t)X * N + Y) * N + Z;
        d.F[0][f] = fxx;
        d.F[1][f] = fxy;
        if (Z == X) d.I[(int64_t)X * N + Y] = fxx;
    }
}

// Table launches enumerate blocks over the correlated trajectories only:
// block blk is entry (blk - tprefix[j]) of the N x table_blocks_per_x(N)
// block space of listed trajectory j.
__device__ inline int table_block(const BatchDev& b, int64_t blk, int& e) {
    int lo = 0, hi = b.ncorr - 1;
    while (lo < hi) {  // last correlated trajectory whose blocks start at or before blk
        const int mid = (lo + hi + 1) >> 1;
        if (b.tprefix[mid] <= blk) lo = mid; else hi = mid - 1;
    }
    e = (int)(blk - b.tprefix[lo]);
    return b.tlist[lo];
}

// Stage set-up of every trajectory in one launch: the G1 part (Y1 = S1 + c K1,
// the G1 pairs, h = h_HF[Y1], W tables) and the two table families, which
// form Y1 and W themselves, so that none of the three waits for another.
// Blocks [0, gA * ntraj): G1 part, gA blocks of SETUP_THREADS elements per
// trajectory; then tblocks Qp blocks and tblocks F blocks (one per X and 64 x 64
// output block).
__global__ __launch_bounds__(SETUP_THREADS) void k_stage_setup(BatchDev b, StageRef r, int herm, int gA) {
    __shared__ StageParams sp;
    __shared__ __align__(16) TileLds<64> L;
    const StageShared ss = stage_of(r);
    const int64_t nA = (int64_t)gA * b.ntraj, blk = blockIdx.x;
    int ti, e, role;
    if (blk < nA) {
        role = 0;
        ti = (int)(blk / gA);
        e = (int)(blk % gA);
    } else {
        role = blk < nA + b.tblocks ? 1 : 2;
        ti = table_block(b, role == 1 ? blk - nA : blk - nA - b.tblocks, e);
    }
    const DevState& d = b.ds[ti];
    const TrajParams& tp = b.tp[ti];
    if (!active(d, tp, ss)) return;
    if (role == 0 && e * (int)blockDim.x >= d.N * d.N) return;
    // built per block in LDS: the V table is indexed at run time and would
    // otherwise live in per-lane scratch
    fill_sp(sp, tp, ss);
    __syncthreads();
    if (role == 0) {
        stage_g1_block(d, ss, sp, e, reinterpret_cast<double2*>(L.xr));
        return;
    }
    const int nb = (d.N + TG - 1) / TG, per_x = table_blocks_per_x(d.N);
    const int X = e / per_x, rb = (e % per_x) / nb, cb = (e % per_x) % nb;
    if (role == 1)
        q_tables_block(d, ss, sp, X, rb, cb, L);
    else
        f_tables_block(d, sp, X, rb, cb, herm != 0, L);
}

// ---- precision engines ------------------------------------------------------
// Each engine owns the LDS image of a tile X and of h, writes a tile element
// converted from FP64, and runs O = h.X - X.h on its MFMA type.  Results come
// back as FP64 (re, im) per 16x16 block in MFMA result order; row_of() maps
// (lane, j) to the block row.

typedef float f4 __attribute__((ext_vector_type(4)));

template <int NP, int NW, int PREC>
struct Engine;

template <int NP, int NW>
struct Engine<NP, NW, G12_PREC_FP64> {
    typedef TileLds<NP> Lds;
    static constexpr int P = Lds::P;
    // output staging of the pair-symmetric pass: the X planes (pitch OP),
    // free once the products have read them
    typedef double Out;
    static constexpr int OP = P;
    __device__ static Out* out_plane(Lds& L, int c) { return c ? L.xi : L.xr; }
    __device__ static void put_h(Lds& L, int m, int n, double2 v) {
        L.hr[m * P + n] = v.x;
        L.hi[m * P + n] = v.y;
    }
    __device__ static void put_x(Lds& L, int m, int n, double2 v) {
        L.xr[m * P + n] = v.x;
        L.xi[m * P + n] = v.y;
    }
    __device__ static void zero(Lds&, int, int) {}
    __device__ static int row_of(int lane, int j) { return 4 * j + (lane >> 4); }
    // mid(add) runs between the h.X and X.h products; add(u, j, v) adds v to
    // result element (u, j) through the engine's accumulators
    template <class Mid>
    __device__ static void compute(const Lds& L, int hmode, d4 (&ar)[TileCfg<NP, NW>::PW],
                                   d4 (&ai)[TileCfg<NP, NW>::PW], int wave, int lane, Mid&& mid) {
        tile_mfma_fp64<NP, NW>(L, hmode, ar, ai, wave, lane, mid);
    }
};

// 3M complex products on FP32 MFMA: with S = re + im (formed in registers
// from the re and im operands, so LDS holds two planes per matrix),
//   P1 = Ar Br, P2 = Ai Bi, P3 = As Bs;  Re = P1 - P2, Im = P3 - P1 - P2.
template <int NP>
struct TileLds32 {
    static constexpr int P = NP + 1;
    float xr[NP * P], xi[NP * P];
    float hr[NP * P], hi[NP * P];
};

template <int NP, int NW>
struct Engine<NP, NW, G12_PREC_FP32_3M> {
    typedef TileLds32<NP> Lds;
    static constexpr int P = Lds::P;
    typedef float Out;  // output staging: the X planes
    static constexpr int OP = P;
    __device__ static Out* out_plane(Lds& L, int c) { return c ? L.xi : L.xr; }
    static constexpr int NT = TileCfg<NP, NW>::NT, TT = TileCfg<NP, NW>::TT, PW = TileCfg<NP, NW>::PW;
    __device__ static void put_h(Lds& L, int m, int n, double2 v) {
        L.hr[m * P + n] = (float)v.x;
        L.hi[m * P + n] = (float)v.y;
    }
    __device__ static void put_x(Lds& L, int m, int n, double2 v) {
        L.xr[m * P + n] = (float)v.x;
        L.xi[m * P + n] = (float)v.y;
    }
    __device__ static void zero(Lds&, int, int) {}
    __device__ static int row_of(int lane, int j) { return 4 * (lane >> 4) + j; }
    template <bool NEG>
    __device__ static void product(const float* ar_, const float* ai_, const float* br_, const float* bi_,
                                   f4 (&p1)[PW], f4 (&p2)[PW], f4 (&p3)[PW], int wave, int lane) {
        constexpr bool FULL = (TT % NW) == 0;
        constexpr int KS = NP / 4;
        const int r = lane & 15, kb = (lane >> 4) * KS;
#pragma unroll 4
        for (int j = 0; j < KS; ++j) {
            const int kk = kb + j;  // spread k: conflict-free LDS reads (see tile_product)
            float a1[PW], a2[PW], a3[PW], b1[PW], b2[PW], b3[PW];
#pragma unroll
            for (int u = 0; u < PW; ++u) {
                const int t = wave + NW * u;
                const bool on = FULL || t < TT;
                const int mm = (t / NT) * 16 + r, nn = (t % NT) * 16 + r;
                a1[u] = on ? ar_[mm * P + kk] : 0.f;
                a2[u] = on ? ai_[mm * P + kk] : 0.f;
                b1[u] = on ? br_[kk * P + nn] : 0.f;
                b2[u] = on ? bi_[kk * P + nn] : 0.f;
                a3[u] = a1[u] + a2[u];
                b3[u] = b1[u] + b2[u];
            }
#pragma unroll
            for (int u = 0; u < PW; ++u) p1[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(NEG ? -a1[u] : a1[u], b1[u], p1[u], 0, 0, 0);
#pragma unroll
            for (int u = 0; u < PW; ++u) p2[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(NEG ? -a2[u] : a2[u], b2[u], p2[u], 0, 0, 0);
#pragma unroll
            for (int u = 0; u < PW; ++u) p3[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(NEG ? -a3[u] : a3[u], b3[u], p3[u], 0, 0, 0);
        }
    }
    template <class Mid>
    __device__ static void compute(const Lds& L, int hmode, d4 (&ar)[PW], d4 (&ai)[PW], int wave, int lane,
                                   Mid&& mid) {
        f4 p1[PW], p2[PW], p3[PW];
#pragma unroll
        for (int u = 0; u < PW; ++u) {
            p1[u] = f4{0.f, 0.f, 0.f, 0.f};
            p2[u] = p1[u];
            p3[u] = p1[u];
        }
        if (hmode & 1) product<false>(L.hr, L.hi, L.xr, L.xi, p1, p2, p3, wave, lane);
        // v enters Re = P1 - P2 through P1 and Im = P3 - P1 - P2 through P3
        mid([&](int u, int j, double2 v) {
            p1[u][j] += (float)v.x;
            p3[u][j] += (float)(v.x + v.y);
        });
        if (hmode & 2) product<true>(L.xr, L.xi, L.hr, L.hi, p1, p2, p3, wave, lane);
#pragma unroll
        for (int u = 0; u < PW; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double q1 = p1[u][j], q2 = p2[u][j], q3 = p3[u][j];
                ar[u][j] = q1 - q2;
                ai[u][j] = q3 - q1 - q2;
            }
    }
};

// BF16 hi/lo split on BF16 MFMA (16x16x32): every operand v = hi + lo with
// hi = bf16(v), lo = bf16(v - hi); a real product keeps hi.hi + hi.lo + lo.hi
// (the lo.lo term is below the split's own rounding).  Four real products per
// complex product, FP32 accumulation.  A operands are stored row-major, B
// operands transposed, so each lane reads 8 consecutive k (16 bytes).
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

template <int NP>
struct TileLdsBf {
    static constexpr int KP = ((NP + 31) / 32) * 32;  // k extent padded to the MFMA depth
    static constexpr int PB = KP + 8;                  // pitch (bf16); 16-byte aligned rows
    // [plane][row][k]; planes: 0-3 h (r hi, r lo, i hi, i lo) row-major,
    // 4-7 h transposed, 8-11 X row-major, 12-15 X transposed
    __bf16 v[16][NP * PB];
};

__device__ __forceinline__ bf8 bf_neg(bf8 x) {
    u4v u = __builtin_bit_cast(u4v, x);
    u ^= (u4v){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    return __builtin_bit_cast(bf8, u);
}

template <int NP, int NW>
struct Engine<NP, NW, G12_PREC_BF16X3> {
    typedef TileLdsBf<NP> Lds;
    static constexpr int PB = Lds::PB, KP = Lds::KP;
    // output staging: two FP32 planes over the X planes (8-15); their zero
    // padding has to be restored afterwards when N < NP
    typedef float Out;
    static constexpr int OP = NP + 1;
    static_assert(2 * NP * OP * sizeof(float) <= 8 * NP * PB * sizeof(__bf16), "staging fits the X planes");
    __device__ static Out* out_plane(Lds& L, int c) { return reinterpret_cast<float*>(&L.v[8][0]) + c * NP * OP; }
    static constexpr int NT = TileCfg<NP, NW>::NT, TT = TileCfg<NP, NW>::TT, PW = TileCfg<NP, NW>::PW;
    __device__ static void split(double v, __bf16& hi, __bf16& lo) {
        hi = (__bf16)(float)v;
        lo = (__bf16)(float)(v - (double)(float)hi);
    }
    __device__ static void put(Lds& L, int base, int m, int n, double2 v) {
        __bf16 rh, rl, ih, il;
        split(v.x, rh, rl);
        split(v.y, ih, il);
        L.v[base + 0][m * PB + n] = rh;
        L.v[base + 1][m * PB + n] = rl;
        L.v[base + 2][m * PB + n] = ih;
        L.v[base + 3][m * PB + n] = il;
        L.v[base + 4][n * PB + m] = rh;  // transposed copy
        L.v[base + 5][n * PB + m] = rl;
        L.v[base + 6][n * PB + m] = ih;
        L.v[base + 7][n * PB + m] = il;
    }
    __device__ static void put_h(Lds& L, int m, int n, double2 v) { put(L, 0, m, n, v); }
    __device__ static void put_x(Lds& L, int m, int n, double2 v) { put(L, 8, m, n, v); }
    __device__ static void zero(Lds& L, int tid, int nthr) {
        unsigned* w = reinterpret_cast<unsigned*>(&L);
        for (int i = tid; i < (int)(sizeof(Lds) / 4); i += nthr) w[i] = 0u;
    }
    __device__ static int row_of(int lane, int j) { return 4 * (lane >> 4) + j; }
    __device__ static bf8 frag(const __bf16* plane, int row, int k) {
        return *reinterpret_cast<const bf8*>(plane + row * PB + k);
    }
    // acc += sign * A.B ; A planes abase (row-major), B planes bbase (transposed)
    template <bool NEG>
    __device__ static void product(const Lds& L, int abase, int bbase, f4 (&cr)[PW], f4 (&ci)[PW], int wave,
                                   int lane) {
        constexpr bool FULL = (TT % NW) == 0;
        const int r = lane & 15, kq = lane >> 4;
#pragma unroll
        for (int k0 = 0; k0 < KP; k0 += 32) {
            const int kk = k0 + 8 * kq;
#pragma unroll
            for (int u = 0; u < PW; ++u) {
                const int t = wave + NW * u;
                if (!(FULL || t < TT)) continue;
                const int mm = (t / NT) * 16 + r, nn = (t % NT) * 16 + r;
                bf8 arh = frag(L.v[abase + 0], mm, kk), arl = frag(L.v[abase + 1], mm, kk);
                bf8 aih = frag(L.v[abase + 2], mm, kk), ail = frag(L.v[abase + 3], mm, kk);
                const bf8 brh = frag(L.v[bbase + 0], nn, kk), brl = frag(L.v[bbase + 1], nn, kk);
                const bf8 bih = frag(L.v[bbase + 2], nn, kk), bil = frag(L.v[bbase + 3], nn, kk);
                if (NEG) {
                    arh = bf_neg(arh);
                    arl = bf_neg(arl);
                    aih = bf_neg(aih);
                    ail = bf_neg(ail);
                }
                const bf8 naih = bf_neg(aih), nail = bf_neg(ail);
                // Re += Ar Br - Ai Bi
                cr[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh, brh, cr[u], 0, 0, 0);
                cr[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh, brl, cr[u], 0, 0, 0);
                cr[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arl, brh, cr[u], 0, 0, 0);
                cr[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(naih, bih, cr[u], 0, 0, 0);
                cr[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(naih, bil, cr[u], 0, 0, 0);
                cr[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nail, bih, cr[u], 0, 0, 0);
                // Im += Ar Bi + Ai Br
                ci[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh, bih, ci[u], 0, 0, 0);
                ci[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh, bil, ci[u], 0, 0, 0);
                ci[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arl, bih, ci[u], 0, 0, 0);
                ci[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih, brh, ci[u], 0, 0, 0);
                ci[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih, brl, ci[u], 0, 0, 0);
                ci[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ail, brh, ci[u], 0, 0, 0);
            }
        }
    }
    template <class Mid>
    __device__ static void compute(const Lds& L, int hmode, d4 (&ar)[PW], d4 (&ai)[PW], int wave, int lane,
                                   Mid&& mid) {
        f4 cr[PW], ci[PW];
#pragma unroll
        for (int u = 0; u < PW; ++u) {
            cr[u] = f4{0.f, 0.f, 0.f, 0.f};
            ci[u] = cr[u];
        }
        if (hmode & 1) product<false>(L, 0, 12, cr, ci, wave, lane);  // + h.X : A = h, B = X^T planes
        mid([&](int u, int j, double2 v) {
            cr[u][j] += (float)v.x;
            ci[u][j] += (float)v.y;
        });
        if (hmode & 2) product<true>(L, 8, 4, cr, ci, wave, lane);    // - X.h : A = X, B = h^T planes
#pragma unroll
        for (int u = 0; u < PW; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                ar[u][j] = cr[u][j];
                ai[u][j] = ci[u][j];
            }
    }
};

// Tile-pass arguments.  The tile list of a launch is the concatenation, over
// the trajectories of one bucket, of nspin * N^2 tiles per trajectory.
struct BTileArgs {
    BatchDev b;
    Bucket k;
    StageRef r;
    int tiling;       // 0: strided (p2,q2) tiles, 1: contiguous (p1,q1) tiles
    int fused;        // 1: fused right-hand side epilogue (contiguous tiling)
    int hmode;        // 1: +h.X, 2: -X.h, 3: both
    int accumulate;   // unfused: R += O instead of R = O
    int nspin;
    int use_y;        // stage input from the Y buffer (after stage 0, or always when compact; 0: always S)
    int compact;      // R and Y held as complex64 (mixed-precision runs)
    int sources;      // runs: each pass adds its share of the G2 sources (see SrcVec)
    int sym;          // pair-symmetric path: contiguous tiles, half of every source (see k_sym_update)
    const double2* h_override;  // routine calls: h to use instead of the trajectory's h
};

// Persistent-tile schedule.  Workgroups are dealt round-robin to the 8 XCDs;
// each XCD owns a contiguous range of the tile list and its workgroups step
// through it together, so tiles that are neighbours in memory (consecutive
// q2 of the strided pass share cache lines) are in flight on one L2 at once.
struct TileSched {
    int64_t total, per_xcd;
    int wg_per_xcd, xcd, local;
    bool xcd_mode;
    __device__ TileSched(int64_t total_, int grid, int b) : total(total_) {
        xcd_mode = (grid % 8) == 0;
        xcd = b & 7;
        local = b >> 3;
        per_xcd = (total + 7) >> 3;  // the last XCD's range may be shorter
        wg_per_xcd = grid >> 3;
    }
    // tile of iteration `it` for this workgroup, or -1 when done
    __device__ int64_t tile(int it, int b, int grid) const {
        if (xcd_mode) {
            const int64_t i = (int64_t)it * wg_per_xcd + local;
            const int64_t g = xcd * per_xcd + i;
            return (i < per_xcd && g < total) ? g : -1;
        }
        const int64_t i = b + (int64_t)it * grid;
        return i < total ? i : -1;
    }
};

__device__ inline StageShared uni_stage(const StageRef& r) {
    const StageShared v = stage_of(r);
    StageShared u;
    u.c = uni_d(v.c);
    u.cn = uni_d(v.cn);
    u.wk = uni_d(v.wk);
    u.dt = uni_d(v.dt);
    u.first = uni(v.first);
    u.last = uni(v.last);
    u.lam = uni_d(v.lam);
    u.phi = uni_d(v.phi);
    u.gate = uni(v.gate);
    u.as_phase = uni(v.as_phase);
    u.step = uni(v.step);
    u.slot = uni(v.slot);
    u.fold = uni(v.fold);
    return u;
}

// Position of global tiles in the bucket: the listed trajectory that owns a
// tile, kept with its tile range so that the tiles of one trajectory are
// located without any memory access.
struct Cursor {
    int64_t begin = 0, end = 0;  // global tiles [begin, end) of the trajectory
    int traj = -1, N = 0;
    bool on = false;             // the trajectory takes part in this stage
};

__device__ inline void seek(const BTileArgs& a, const StageShared& ss, Cursor& c, int64_t gt) {
    if (gt >= c.begin && gt < c.end) return;
    int lo = 0, hi = a.k.n - 1;
    while (lo < hi) {  // last listed trajectory with prefix <= gt
        const int mid = (lo + hi + 1) >> 1;
        if (a.k.prefix[mid] <= gt) lo = mid; else hi = mid - 1;
    }
    c.begin = uni64(a.k.prefix[lo]);
    c.end = uni64(a.k.prefix[lo + 1]);
    c.traj = uni(a.k.list[lo]);
    c.N = uni(a.b.tp[c.traj].N);
    c.on = uni(active(a.b.ds[c.traj], a.b.tp[c.traj], ss) ? 1 : 0) != 0;
}

// Tile of a trajectory: spin (0 xxxx, 1 xyxy) and index within the spin;
// ok = false for the unused slots of the pair-symmetric list.
struct TileRef {
    int traj, N, spin, tile;
    bool ok;
};

// HALF: the pair-symmetric pass's list of tiles (t0, t1) with t0 <= t1, in
// groups of 4 x 4: slot 16 g + 4 i + j of a spin is t0 = 4 A + i, t1 = 4 B + j
// for the group pair g = B (B + 1) / 2 + A (A <= B); slots with t0 > t1 or
// past N are unused.  Neighbouring workgroups then share the rows of the G1
// planes and source tables their tiles read.  16 hgroups(N) slots per spin.
template <bool HALF = false>
__device__ inline TileRef ref_of(const Cursor& c, int64_t gt) {
    TileRef t;
    t.traj = c.traj;
    t.N = c.N;
    t.ok = true;
    if constexpr (HALF) {
        const int H = 16 * hgroups(c.N);
        int local = (int)(gt - c.begin);  // < 2 H
        t.spin = local >= H ? 1 : 0;
        local -= t.spin * H;
        const int g = local >> 4, w = local & 15;
        int B = (int)((sqrt(8.0 * (double)g + 1.0) - 1.0) * 0.5);
        while ((B + 1) * (B + 2) / 2 <= g) ++B;
        while (B * (B + 1) / 2 > g) --B;
        const int t0 = 4 * (g - B * (B + 1) / 2) + (w >> 2), t1 = 4 * B + (w & 3);
        t.ok = t0 <= t1 && t1 < c.N;
        t.tile = t.ok ? t0 * c.N + t1 : 0;
        return t;
    }
    const int local = (int)(gt - c.begin);  // < 2 N^2
    t.spin = local >= c.N * c.N ? 1 : 0;
    t.tile = local - t.spin * c.N * c.N;
    return t;
}

// Buffers of the (trajectory, spin) whose tile is being processed; reloaded
// only when that pair changes.
struct View {
    int traj = -1, spin = -1;
    const double2* X;   // stage input: S at stage 0, else Y = S + c K_prev
    bool xc;            // X held compact
    double2* S;         // stage state (fused: read for Y, updated at the last stage)
    double2* R;         // unfused output / strided-pass result (compact in mixed runs)
    double2* D;         // fused: partial traces of the next stage input
    double2 *A, *Y;     // fused: RK accumulator, next stage input (compact in mixed runs)
    double2* Sc;        // fused, mixed runs: complex64 copy of the new state (last stage)
    const double2* G1;  // the stage G1 g (Y1)
    const double2* G1t; // its conjugate transpose: conj g[Y][X] at [X][Y]
    const double2 *Qp, *F;  // source tables of the spin
    const double* W;        // W_xx (spin 0) or W_xy (spin 1)
};

// Mixed-precision runs keep the stage input compact at every stage: at stage 0
// it is Sc, the complex64 copy of S (written with every new state by the last
// stage's update, or at the start of a run).
// FP64 runs on the pair-symmetric path keep the stage-1 and stage-2 inputs
// apart: the stage-2 input lives in the A buffer (k_sym_update).
__device__ inline const double2* stage_input(const DevState& d, const BTileArgs& a, const StageShared& ss, int spin) {
    if (a.use_y && a.compact) return ss.first ? d.Sc[spin] : d.Y2[spin];
    if (a.use_y && a.sym && ss.fold) return d.A2[spin];
    return (a.use_y && !ss.first) ? d.Y2[spin] : d.S2[spin];
}
// the stage input is compact when it is the Y buffer of a mixed-precision run
__device__ inline bool input_compact(const BTileArgs& a, const StageShared&) {
    return a.compact && a.use_y;
}

template <bool FUSED>
__device__ inline void load_view(View& v, const BTileArgs& a, const StageShared& ss, int traj, int spin) {
    const DevState& d = a.b.ds[traj];
    v.traj = traj;
    v.spin = spin;
    v.X = uni_ptr(stage_input(d, a, ss, spin));
    v.xc = input_compact(a, ss);
    v.S = uni_ptr(d.S2[spin]);
    v.R = uni_ptr(d.R2[spin]);
    v.G1 = uni_ptr(d.Y1);
    v.G1t = uni_ptr(d.G1t);
    v.Qp = uni_ptr(d.Qp[spin]);
    v.F = uni_ptr(d.F[spin]);
    v.W = uni_ptr(spin == 0 ? d.Wxx : d.Wxy);
    if constexpr (FUSED) {
        v.A = uni_ptr(d.A2[spin]);
        v.Y = uni_ptr(d.Y2[spin]);
        v.Sc = uni_ptr(d.Sc[spin]);
        v.D = uni_ptr(d.D[spin]);
    }
}

// Element offsets of a tile within one tensor (32-bit: N <= 64).  Element
// (m, n) of a tile lies at base + m sm + n B in the natural layout
// T[p1][p2][q1][q2] (the FP64 buffers: state S, RK accumulator A, and Y, R of
// FP64 runs), and at base + m sm + (n >> 2) A + (n & 3) B in the blocked
// layout of the complex64 buffers of mixed-precision runs (stage input Y,
// strided-pass result R).  Blocked: each (p1, p2) slice is an Nq x Nq array
// (Nq = N rounded up to a multiple of 4) stored as 4 x 4 blocks over
// (q1, q2), one 128-byte line each, q2 fastest inside a block:
//   (p1 N + p2) Nq^2 + (q1 >> 2) 4 Nq + (q2 >> 2) 16 + (q1 & 3) 4 + (q2 & 3).
// A strided tile (fixed p2, q2; rows p1, columns q1) then reads and writes
// four lanes per line (32-byte runs at a 32-byte pitch) instead of one lane
// per line, and a contiguous tile (fixed p1, q1; rows p2, columns q2) 32-byte
// runs of four q2.  These layouts serve the general path; the pair-symmetric
// path keeps every buffer in the Hermitian half layout (half_geom, hblock).
// Which complex64 buffers use the blocked layout (mixed-precision runs).
#ifndef G12_BLOCKED_Y
#define G12_BLOCKED_Y 1
#endif
#ifndef G12_BLOCKED_R
#define G12_BLOCKED_R 1
#endif
constexpr bool BLK_Y = G12_BLOCKED_Y != 0, BLK_R = G12_BLOCKED_R != 0;

struct TileGeom {
    uint32_t base, sm, A, B;
    int t0, t1;  // the tile's fixed pair: (p2, q2) strided, (p1, q1) contiguous
    __device__ __forceinline__ uint32_t nat(int m, int n) const { return base + (uint32_t)m * sm + (uint32_t)n * B; }
    __device__ __forceinline__ uint32_t blk(int m, int n) const {
        return base + (uint32_t)m * sm + (uint32_t)(n >> 2) * A + (uint32_t)(n & 3) * B;
    }
    template <bool BLK>
    __device__ __forceinline__ uint32_t at(int m, int n) const { return BLK ? blk(m, n) : nat(m, n); }
};

__device__ __forceinline__ TileGeom tile_geom(int tiling, int N, int tile, bool blocked) {
    TileGeom g;
    const uint32_t n = (uint32_t)N;
    g.t0 = tile / N;
    g.t1 = tile % N;
    const uint32_t t0 = (uint32_t)g.t0, t1 = (uint32_t)g.t1;
    if (!blocked) {
        const uint32_t N2 = n * n, N3 = N2 * n;
        if (tiling == 0) {  // strided: fixed (p2, q2) = (t0, t1); m = p1, n = q1
            g.base = t0 * N2 + t1;
            g.sm = N3;
            g.B = n;
        } else {  // contiguous: fixed (p1, q1) = (t0, t1); m = p2, n = q2
            g.base = t0 * N3 + t1 * n;
            g.sm = N2;
            g.B = 1;
        }
        g.A = 4 * g.B;
    } else {
        const uint32_t nq = (n + 3) & ~3u, Q2 = nq * nq;
        if (tiling == 0) {
            g.base = t0 * Q2 + (t1 >> 2) * 16 + (t1 & 3);
            g.sm = n * Q2;
            g.A = 4 * nq;
            g.B = 4;
        } else {
            g.base = t0 * n * Q2 + (t1 >> 2) * 4 * nq + (t1 & 3) * 4;
            g.sm = Q2;
            g.A = 16;
            g.B = 1;
        }
    }
    return g;
}

// Contiguous tile (t0, t1), t0 <= t1, in the Hermitian half layout of the
// pair-symmetric path (hblock): rows p2 at a pitch of N, columns q2
// contiguous, for complex128 and complex64 buffers alike (blk = nat).
__device__ __forceinline__ TileGeom half_geom(int N, int tile) {
    TileGeom g;
    g.t0 = tile / N;
    g.t1 = tile % N;
    g.base = hblock(g.t0, g.t1) * (uint32_t)(N * N);
    g.sm = (uint32_t)N;
    g.B = 1;
    g.A = 4;
    return g;
}

// Stage input of one tile into registers (TILING: 0 strided, 1 contiguous,
// -1 given at run time; HALF: the Hermitian half layout).  Plain loads:
// nothing waits for them before the MFMA loop.
template <int NP, int NW, int TILING, bool FULL, bool HALF = false>
__device__ __forceinline__ void tile_prefetch(const BTileArgs& a, const TileRef& t, const double2* X, bool xc,
                                              double2 (&pre)[TileCfg<NP, NW>::PF]) {
    constexpr int NTHR = TileCfg<NP, NW>::NTHR, PF = TileCfg<NP, NW>::PF;
    // every slot holds a tile element only when the threads divide NP^2
    constexpr bool EXACT = FULL && (NP * NP) % NTHR == 0;
    const int N = FULL ? NP : t.N;
    const int tiling = TILING < 0 ? a.tiling : TILING;
    // one branch on the storage form, outside the unrolled loads
    if (xc) {
        const TileGeom g = HALF ? half_geom(N, t.tile) : tile_geom(tiling, N, t.tile, BLK_Y);
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int e = threadIdx.x + NTHR * i;
            double2 y = make_double2(0.0, 0.0);
            if (EXACT || e < N * N) y = cld(X, g.at<BLK_Y>(e / N, e % N), true);
            pre[i] = y;
        }
    } else {
        const TileGeom g = HALF ? half_geom(N, t.tile) : tile_geom(tiling, N, t.tile, false);
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int e = threadIdx.x + NTHR * i;
            double2 y = make_double2(0.0, 0.0);
            if (EXACT || e < N * N) y = gld(X, g.nat(e / N, e % N));
            pre[i] = y;
        }
    }
}

// The same into raw complex64 registers (mixed-precision engines, whose tile
// products take FP32 or BF16 operands): a compact stage input is loaded as it
// is stored, with nothing waiting for it; a complex128 input (routine-level
// calls only) is rounded on arrival.
template <int NP, int NW, int TILING, bool FULL, bool HALF = false>
__device__ __forceinline__ void tile_prefetch(const BTileArgs& a, const TileRef& t, const double2* X, bool xc,
                                              fv2 (&pre)[TileCfg<NP, NW>::PF]) {
    constexpr int NTHR = TileCfg<NP, NW>::NTHR, PF = TileCfg<NP, NW>::PF;
    // every slot holds a tile element only when the threads divide NP^2
    constexpr bool EXACT = FULL && (NP * NP) % NTHR == 0;
    const int N = FULL ? NP : t.N;
    const TileGeom g = HALF ? half_geom(N, t.tile) : tile_geom(TILING < 0 ? a.tiling : TILING, N, t.tile, xc && BLK_Y);
#pragma unroll
    for (int i = 0; i < PF; ++i) {
        const int e = threadIdx.x + NTHR * i;
        fv2 y = fv2{0.f, 0.f};
        if (EXACT || e < N * N) {
            if (xc) {
                y = *(const GLOBAL fv2*)((const GLOBAL char*)X + (g.at<BLK_Y>(e / N, e % N) << 3));
            } else {
                const double2 d = gld(X, g.nat(e / N, e % N));
                y = fv2{(float)d.x, (float)d.y};
            }
        }
        pre[i] = y;
    }
}

__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(fv2 v) { return make_double2(v.x, v.y); }

// Sources of the G2 equation (psi.h:26-277, pi.h:25-95, G1_G2_t_step.h:55-86
// through the partial-trace tables), for element (p1, p2, q1, q2) of spin s
// with g the stage G1, g11 = g(p1,q1), g22 = g(p2,q2), W = W_xx (s = 0) or
// W_xy (s = 1), Qp[X][Y][Z] = Q[Z][X][Y] and F the tables of the stage:
//   Psi  = g11 g22 (W[p1][p2] - W[q1][q2]) + i g11 (Qp[p2][q2][p1] - Qp[p2][q2][q1])
//          + i g22 (Qp[p1][q1][p2] - Qp[p1][q1][q2])
//   Pi   = g22 (F[p1][q1][p2] - F[p1][q1][q2])
//   Pi^+ = conj( g(q1,p1) (F[q2][p2][q1] - F[q2][p2][p1]) )
// Each tile pass adds the terms that are differences of two vectors over its
// own tile's rows and columns, so that no source needs a matrix per tile:
//   strided tile, fixed (p2, q2), element (m, n) = (p1, q1):
//     g11 (a[m] - b[n]) - conj(g(q1,p1)) (c[n] - c[m]),
//     a[x] = g22 W[x][p2] + i Qp[p2][q2][x], b[x] = g22 W[x][q2] + i Qp[p2][q2][x],
//     c[x] = conj F[q2][p2][x];
//   fused tile, fixed (p1, q1), element (m, n) = (p2, q2):
//     g22 (e[m] - e[n]),  e[x] = i Qp[p1][q1][x] + F[p1][q1][x].
// g11 and conj g(q1,p1) (strided) and g22 (fused) are read per element from
// the stage G1 and its conjugate transpose (two N x N planes, contiguous along
// a tile's columns), which every tile shares; the vectors are loaded by one
// wave, one element per lane (N <= 64), and reach the lanes that need them
// through LDS.  The tables depend only on the stage G1 and on the partial
// traces of the stage input, so the strided pass can run after them.
struct SrcVec {
    double2 q, f, g;  // Qp row, F row, g22 (strided only)
    double w0, w1;    // W rows at p2 and q2 (strided only)
};

// Source vectors of tile t from the buffers of its (trajectory, spin) in v.
// SYM (pair-symmetric path, contiguous tiles): the fused pass's F row and
// the W rows and g11 of the strided form.
template <bool FUSED, bool SYM = false>
__device__ __forceinline__ void src_load(const View& v, const TileRef& t, int lane, SrcVec& sv) {
    const int N = t.N;
    const int t0 = t.tile / N, t1 = t.tile % N;
    const uint32_t n = (uint32_t)N;
    const uint32_t x = lane < N ? lane : 0;  // lanes past N read a valid element they never use
    sv.q = gld(v.Qp, ((uint32_t)t0 * n + t1) * n + x);
    if constexpr (FUSED) {
        sv.f = gld(v.F, ((uint32_t)t0 * n + t1) * n + x);
    } else if constexpr (SYM) {
        sv.f = gld(v.F, ((uint32_t)t0 * n + t1) * n + x);
        sv.w0 = gld1(v.W, (uint32_t)t0 * n + x);
        sv.w1 = gld1(v.W, (uint32_t)t1 * n + x);
        sv.g = gld(v.G1, (uint32_t)t0 * n + t1);
    } else {
        sv.f = gld(v.F, ((uint32_t)t1 * n + t0) * n + x);
        sv.w0 = gld1(v.W, (uint32_t)t0 * n + x);  // W is symmetric
        sv.w1 = gld1(v.W, (uint32_t)t1 * n + x);
        sv.g = gld(v.G1, (uint32_t)t0 * n + t1);
    }
}

__device__ __forceinline__ double2 lane_value(double2 v, int src) {
    return make_double2(__shfl(v.x, src), __shfl(v.y, src));
}

// FULL: every tile of the launch has N = NP, so the element bounds checks and
// the divisions by N drop out at compile time.
// SYM (unfused only): the product pass of the pair-symmetric path.
// Contiguous tiles X[p2][q2] at fixed (p1, q1) = (t0, t1); the output
// R = h.X - X.h + g22 (ea[p2] - eb[q2]) with e = i Qp[t0][t1][.] +
// F[t0][t1][.] and ea = e + (g11/2) W[t0][.], eb = e + (g11/2) W[t1][.]: the
// fused pass's terms and half of the Psi term g11 g22 (W[p1][p2] -
// W[q1][q2]); the update kernel adds the pair-exchanged R (k_sym_update).
// The state is Hermitian as well, so R is anti-Hermitian as an N^2 x N^2
// matrix: the tile (t1, t0) of R is -(tile (t0, t1))^dagger, because
// (h X - X h)^dagger = X^dagger h - h X^dagger with the input tile
// (t1, t0) = (tile (t0, t1))^dagger, Qp[t1][t0][.] = conj Qp[t0][t1][.] and
// F[t1][t0][.] = -conj F[t0][t1][.] for G1 anti-Hermitian and D Hermitian.
// Only the tiles t0 <= t1 are computed and stored, in the Hermitian half
// layout (half_geom): off-diagonal tiles straight from the MFMA registers,
// diagonal tiles through LDS as (R - R^dagger) / 2.
template <int NP, int NW, int PREC, bool FUSED, bool FULL, bool SYM = false>
__global__ __launch_bounds__(64 * NW) void k_tile(BTileArgs a) {
    static_assert(!(FUSED && SYM), "the pair-symmetric products run unfused");
    typedef Engine<NP, NW, PREC> E;
    __shared__ typename E::Lds L;
    // source vectors of the current tile, one element per row/column index:
    // strided (a, b, c), fused (e).  The FP64 fused pass keeps them in
    // registers and permutes them across lanes instead (measured 4% faster
    // there; through LDS everywhere else)
    constexpr bool SRC_LDS = !(FUSED && PREC == G12_PREC_FP64);
    __shared__ double2 srcv[FUSED ? 1 : 3][NP];
    constexpr int NT = TileCfg<NP, NW>::NT, TT = TileCfg<NP, NW>::TT, PW = TileCfg<NP, NW>::PW;
    constexpr int NTHR = TileCfg<NP, NW>::NTHR, PF = TileCfg<NP, NW>::PF;
    constexpr int TILING = (FUSED || SYM) ? 1 : -1;
    constexpr bool ALLB = (TT % NW) == 0;                 // every wave owns PW output blocks
    constexpr bool EXACT = FULL && (NP * NP) % NTHR == 0;  // every prefetch slot holds a tile element
    const StageShared ss = uni_stage(a.r);
    const TileSched sched(a.k.tiles, gridDim.x, blockIdx.x);
    E::zero(L, threadIdx.x, NTHR);
    // next tile of a trajectory that takes part in this stage, or -1
    Cursor c;
    auto next_tile = [&](int& it) -> int64_t {
        for (;;) {
            const int64_t g = sched.tile(it, blockIdx.x, gridDim.x);
            if (g < 0) return -1;
            seek(a, ss, c, g);
            if (c.on && (!SYM || ref_of<SYM>(c, g).ok)) return g;
            ++it;
        }
    };
    View v;
    int it = 0;
    int64_t gt = next_tile(it);
    if (gt < 0) return;
    TileRef tr = ref_of<SYM>(c, gt);
    // Stage-input prefetch: a tile's X loads into registers while earlier
    // tiles are multiplied.  The mixed-precision engines hold raw complex64
    // elements (their stage input is compact), so two tiles fit in the
    // registers of one complex128 tile and the tile after next is in flight
    // too: their shorter MFMA loops do not cover one tile's load latency.
    constexpr bool DEEP = PREC != G12_PREC_FP64;
    typedef typename std::conditional<DEEP, fv2, double2>::type PreT;
    PreT P0[PF], P1[PF];
    auto prefetch = [&](const TileRef& t, PreT (&P)[PF]) {
        const double2* X = (t.traj == v.traj && t.spin == v.spin)
                               ? v.X
                               : uni_ptr(stage_input(a.b.ds[t.traj], a, ss, t.spin));
        tile_prefetch<NP, NW, TILING, FULL, SYM>(a, t, X, input_compact(a, ss), P);
    };
    prefetch(tr, P0);
    int64_t g1 = -1;  // DEEP: the tile after tr, in flight in the other buffer
    TileRef t1r = tr;
    if constexpr (DEEP) {
        ++it;
        g1 = next_tile(it);
        if (g1 >= 0) {
            t1r = ref_of<SYM>(c, g1);
            prefetch(t1r, P1);
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // source vectors of the current tile: loaded after the previous tile's
    // epilogue, used in this tile's
    // the first wave loads them (one element per lane) and writes them to LDS
    // with the tile's stage input
    const bool lead = uni(threadIdx.x >> 6) == 0 || !SRC_LDS;
    SrcVec sv{};
    if (a.sources && lead) {
        View v0;
        load_view<FUSED>(v0, a, ss, tr.traj, tr.spin);
        src_load<FUSED, SYM>(v0, tr, lane, sv);
    }
    // one tile; Pc holds its stage input and is refilled with the tile after
    // the ones already in flight
    auto step = [&](PreT (&Pc)[PF]) -> bool {
        if (gt < 0) return false;
        const int N = FULL ? NP : tr.N;
        // element (m, n) of output block t lies inside the tile
        auto inside = [&](int t, int m, int n) {
            return (ALLB || t < TT) && (FULL || (m < N && n < N));
        };
        lds_sync();  // previous tile's readers of X (and h) are done
        if (tr.traj != v.traj || tr.spin != v.spin) {
            const bool new_h = tr.traj != v.traj;
            load_view<FUSED>(v, a, ss, tr.traj, tr.spin);
            if (new_h) {  // h (and the zero padding of X) for this trajectory
                const double2* h = a.h_override ? a.h_override : a.b.ds[tr.traj].h;
                for (int e = threadIdx.x; e < NP * NP; e += NTHR) {
                    const int m = e / NP, n = e % NP;
                    double2 hv = make_double2(0.0, 0.0);
                    if (m < N && n < N) hv = h[m * N + n];
                    E::put_h(L, m, n, hv);
                    E::put_x(L, m, n, make_double2(0.0, 0.0));
                }
                lds_sync();
            }
        }
        // natural geometry (S, A; Y and R of FP64 runs) and the blocked one of
        // the complex64 buffers (Y, R) of mixed-precision runs
        const int tiling = TILING < 0 ? a.tiling : TILING;
        const TileGeom gn = tile_geom(tiling, N, tr.tile, false);
        const TileGeom gc = tile_geom(tiling, N, tr.tile, true);
        const int t0 = gn.t0, t1 = gn.t1;
#pragma unroll
        for (int i = 0; i < PF; ++i) {
            const int e = threadIdx.x + NTHR * i;
            if (EXACT || e < N * N) E::put_x(L, e / N, e % N, to_d2(Pc[i]));
        }
        if (SRC_LDS && a.sources && lead && lane < N) {
            if constexpr (FUSED) {
                srcv[0][lane] = cadd(cmul_i(sv.q), sv.f);
            } else if constexpr (SYM) {
                const double2 e = cadd(cmul_i(sv.q), sv.f), hg = cscale(sv.g, 0.5);
                srcv[0][lane] = cadd(e, cscale(hg, sv.w0));
                srcv[1][lane] = cadd(e, cscale(hg, sv.w1));
            } else {
                srcv[0][lane] = cadd(cscale(sv.g, sv.w0), cmul_i(sv.q));
                srcv[1][lane] = cadd(cscale(sv.g, sv.w1), cmul_i(sv.q));
                srcv[2][lane] = cconj(sv.f);
            }
        }
        lds_sync();
        ++it;
        const int64_t gnew = next_tile(it);
        TileRef rnew = tr;
        if (gnew >= 0) {  // its loads overlap the MFMA loop(s) before its turn
            rnew = ref_of<SYM>(c, gnew);
            prefetch(rnew, Pc);
        }
        int64_t gnext = gnew;  // the tile after tr
        TileRef nt = rnew;
        if constexpr (DEEP) {
            gnext = g1;
            nt = t1r;
            g1 = gnew;
            t1r = rnew;
        }

        // fused pass: this lane's R and S elements load while the MFMA loop
        // runs.  Between the two products R is added to the accumulators and
        // the RK accumulator A starts loading into the registers R held, so
        // A's latency is hidden without more registers.
        double2 rpre[FUSED ? PW : 1][4], spre[FUSED ? PW : 1][4], apre[FUSED ? PW : 1][4];
        if constexpr (FUSED) {
            auto preload = [&](auto compact) {
#pragma unroll
                for (int u = 0; u < PW; ++u) {
                    const int t = wave + NW * u;
                    const int n = (t % NT) * 16 + (lane & 15);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int m = (t / NT) * 16 + E::row_of(lane, j);
                        rpre[u][j] = spre[u][j] = make_double2(0.0, 0.0);
                        if (inside(t, m, n)) {
                            constexpr bool C = decltype(compact)::value;
                            rpre[u][j] = cld(v.R, (C && BLK_R) ? gc.blk(m, n) : gn.nat(m, n), C);
                            spre[u][j] = gld(v.S, gn.nat(m, n));
                        }
                    }
                }
            };
            if (a.compact)
                preload(std::true_type{});
            else
                preload(std::false_type{});
        }

        d4 ar[PW], ai[PW];
        E::compute(L, a.hmode, ar, ai, wave, lane, [&](auto&& add) {
            if constexpr (FUSED) {
#pragma unroll
                for (int u = 0; u < PW; ++u) {
                    const int t = wave + NW * u;
                    const int n = (t % NT) * 16 + (lane & 15);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int m = (t / NT) * 16 + E::row_of(lane, j);
                        add(u, j, rpre[u][j]);
                        apre[u][j] = make_double2(0.0, 0.0);
                        if (!ss.first && inside(t, m, n)) apre[u][j] = gld(v.A, gn.nat(m, n));
                    }
                }
            }
        });

        if constexpr (SYM) {
          // the tile (t0, t1) of R in the Hermitian half layout (its mirror
          // (t1, t0) = -(.)^dagger is not stored).  t0 != t1: straight from
          // the MFMA layout (256-byte complex128 or 128-byte complex64 rows);
          // t0 = t1 through LDS, which forms (R - R^dagger) / 2
          if (t0 != t1) {
            const TileGeom ga = half_geom(N, t0 * N + t1);
            // two forms, with and without non-temporal stores, chosen per tile
            // (a run-time choice per store would be merged and lose the hint)
            auto direct = [&](auto ntc) {
                constexpr bool NTS = decltype(ntc)::value;
#pragma unroll
                for (int u = 0; u < PW; ++u) {
                    const int t = wave + NW * u;
                    if (!ALLB && t >= TT) continue;
                    const int n = (t % NT) * 16 + (lane & 15);
                    const double2 bn = a.sources ? srcv[1][n] : double2{};
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int m = (t / NT) * 16 + E::row_of(lane, j);
                        if (!inside(t, m, n)) continue;
                        double2 val = make_double2(ar[u][j], ai[u][j]);
                        if (a.sources)  // g22 (ea[p2] - eb[q2])
                            val = cadd(val, cmul(gld(v.G1, (uint32_t)(m * N + n)), csub(srcv[0][m], bn)));
                        if constexpr (NTS)
                            cst_nt(v.R, ga.nat(m, n), val, a.compact);
                        else
                            cst(v.R, ga.nat(m, n), val, a.compact);
                    }
                }
            };
            if (use_nt(a.compact, N))
                direct(std::true_type{});
            else
                direct(std::false_type{});
          } else {
            // stage R (products + sources) in LDS once every wave has read X
            typedef typename E::Out OT;
            OT* ore = E::out_plane(L, 0);
            OT* oim = E::out_plane(L, 1);
            constexpr int OP = E::OP;
            lds_sync();
#pragma unroll
            for (int u = 0; u < PW; ++u) {
                const int t = wave + NW * u;
                if (!ALLB && t >= TT) continue;
                const int n = (t % NT) * 16 + (lane & 15);
                const double2 bn = a.sources ? srcv[1][n] : double2{};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int m = (t / NT) * 16 + E::row_of(lane, j);
                    if (!inside(t, m, n)) continue;
                    double2 val = make_double2(ar[u][j], ai[u][j]);
                    if (a.sources)  // g22 (ea[p2] - eb[q2])
                        val = cadd(val, cmul(gld(v.G1, (uint32_t)(m * N + n)), csub(srcv[0][m], bn)));
                    ore[m * OP + n] = (OT)val.x;
                    oim[m * OP + n] = (OT)val.y;
                }
            }
            lds_sync();
            // rows of tile (t0, t1); t0 = t1: its anti-Hermitian part (R - R^dagger) / 2
            auto write_out = [&](auto compact, auto ntc) {
                constexpr bool C = decltype(compact)::value, NTS = decltype(ntc)::value;
                const TileGeom ga = half_geom(N, t0 * N + t1);
#pragma unroll
                for (int i = 0; i < PF; ++i) {
                    const int e = threadIdx.x + NTHR * i;
                    if (!(EXACT || e < N * N)) continue;
                    const int m = e / N, n = e % N;
                    const double vr = ore[m * OP + n], vi = oim[m * OP + n];
                    if (t0 == t1) {
                        const double wr = ore[n * OP + m], wi = oim[n * OP + m];
                        cst(v.R, ga.nat(m, n), make_double2((vr - wr) * 0.5, (vi + wi) * 0.5), C);
                    } else {
                        if constexpr (NTS)
                            cst_nt(v.R, ga.nat(m, n), make_double2(vr, vi), C);
                        else
                            cst(v.R, ga.nat(m, n), make_double2(vr, vi), C);
                    }
                }
            };
            if (a.compact) {
                if (use_nt(true, N))
                    write_out(std::true_type{}, std::true_type{});
                else
                    write_out(std::true_type{}, std::false_type{});
            } else {
                write_out(std::false_type{}, std::true_type{});
            }
            if constexpr (PREC == G12_PREC_BF16X3) {
                // the staging covered the zero padding of the bf16 X planes
                // (k runs to the MFMA depth KP >= NP)
                if (N < TileLdsBf<NP>::KP) {
                    lds_sync();
                    unsigned* w = reinterpret_cast<unsigned*>(ore);
                    for (int i = threadIdx.x; i < 2 * NP * OP; i += NTHR) w[i] = 0u;
                }
            }
          }
        } else if constexpr (!FUSED) {
            double2* out = v.R;
            // strided-pass sources: g11 (va[m] - vb[n]) - conj g(q1,p1) (vc[n] - vc[m]),
            // va, vb, vc in srcv[0..2]
            auto store = [&](auto compact) {
                constexpr bool C = decltype(compact)::value;
#pragma unroll
                for (int u = 0; u < PW; ++u) {
                    const int t = wave + NW * u;
                    if (!ALLB && t >= TT) continue;
                    const int n = (t % NT) * 16 + (lane & 15);
                    double2 bn{}, cn{};
                    if (a.sources) {
                        bn = srcv[1][n];
                        if constexpr (!SYM) cn = srcv[2][n];
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int m = (t / NT) * 16 + E::row_of(lane, j);
                        double2 am{}, cm{};
                        if (a.sources) {
                            am = srcv[0][m];
                            if constexpr (!SYM) cm = srcv[2][m];
                        }
                        if (inside(t, m, n)) {
                            const uint32_t off = (C && BLK_R) ? gc.blk(m, n) : gn.nat(m, n);
                            double2 val = make_double2(ar[u][j], ai[u][j]);
                            if (a.sources) {
                                const uint32_t gi = (uint32_t)(m * N + n);
                                if constexpr (SYM) {  // g22 (ea[p2] - eb[q2])
                                    val = cadd(val, cmul(gld(v.G1, gi), csub(am, bn)));
                                } else {
                                    const double2 g11 = gld(v.G1, gi), gt = gld(v.G1t, gi);
                                    val = cadd(val, csub(cmul(g11, csub(am, bn)), cmul(gt, csub(cn, cm))));
                                }
                            }
                            if (a.accumulate) val = cadd(cld(out, off, C), val);
                            cst(out, off, val, C);
                        }
                    }
                }
            };
            if (a.compact)
                store(std::true_type{});
            else
                store(std::false_type{});
        } else {
            const double2 ve = cadd(cmul_i(sv.q), sv.f);  // e = i Qp + F (register form)
            // fused update for tile (p1, q1) = (t0, t1), element (p2, q2) = (m, n):
            // the accumulators hold R + O (R: the strided-pass terms and their
            // sources); with this pass's sources g22 (ve[m] - ve[n]),
            // K = -i (R + O + sources); then A += w K and the next stage input
            // Y = S + c_next K, or at the last stage S += dt (A + w K) / 6.
            auto update = [&](auto compact) {
                constexpr bool C = decltype(compact)::value;
#pragma unroll
                for (int u = 0; u < PW; ++u) {
                    const int t = wave + NW * u;
                    if (!ALLB && t >= TT) continue;
                    const int n = (t % NT) * 16 + (lane & 15);
                    const double2 en = !a.sources ? double2{} : SRC_LDS ? srcv[0][n] : lane_value(ve, n);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int m = (t / NT) * 16 + E::row_of(lane, j);
                        const double2 em = !a.sources ? double2{} : SRC_LDS ? srcv[0][m] : lane_value(ve, m);
                        if (!inside(t, m, n)) continue;
                        const uint32_t off = gn.nat(m, n);                        // S, A
                        const uint32_t offc = (C && BLK_Y) ? gc.blk(m, n) : off;  // Y
                        double2 rr = make_double2(ar[u][j], ai[u][j]);  // R + O
                        if (a.sources) rr = cadd(rr, cmul(gld(v.G1, (uint32_t)(m * N + n)), csub(em, en)));
                        const double2 K = cmul_mi(rr);
                        const double2 wK = cscale(K, ss.wk);
                        const double2 s = spre[u][j];
                        double2 next;  // input of the next stage (the new state after the last)
                        if (!ss.last) {
                            if (ss.first)
                                gst(v.A, off, wK);
                            else
                                gst(v.A, off, cadd(apre[u][j], wK));
                            next = cadd(s, cscale(K, ss.cn));
                            cst(v.Y, offc, next, C);
                        } else {
                            const double2 A = cadd(apre[u][j], wK);
                            next = cadd(s, cdiv6(cscale(A, ss.dt)));
                            gst(v.S, off, next);
                            if constexpr (C) cst(v.Sc, offc, next, true);  // next step's stage-0 input
                        }
                        // partial trace of the next stage input, FP64:
                        // D[X=p1][Y=q1][Z=p2] for p2 = q2
                        if (m == n) gst(v.D, (uint32_t)((t0 * N + t1) * N + m), next);
                    }
                }
            };
            if (a.compact)
                update(std::true_type{});
            else
                update(std::false_type{});
        }
        if (a.sources && lead && gnext >= 0) {
            if (nt.traj == v.traj && nt.spin == v.spin) {
                src_load<FUSED, SYM>(v, nt, lane, sv);
            } else {
                View vn;
                load_view<FUSED>(vn, a, ss, nt.traj, nt.spin);
                src_load<FUSED, SYM>(vn, nt, lane, sv);
            }
        }
        gt = gnext;
        tr = nt;
        return true;
    };
    if constexpr (DEEP) {
        while (step(P0) && step(P1)) {
        }
    } else {
        while (step(P0)) {
        }
    }
}

// ---------------------------------------------------------------------------
// Wide tiles: 64 < N <= 128, FP64.  An N x N complex128 tile (up to 256 KB)
// and h do not fit in LDS together, so both products stream their operands
// through LDS in k-chunks of WK: h.X stages columns [k0, k0 + WK) of h and
// rows [k0, k0 + WK) of X, then -X.h columns of X (negated as they are
// staged, so both products accumulate) and rows of h.  The next chunk loads
// into registers while the current one is multiplied.  The products are 3M on
// v_mfma_f64_16x16x4f64 as in k_tile, and the epilogues (sources, RK update,
// partial traces) are those of k_tile's FP64 passes.  The (N/16)^2 output
// blocks of 16x16 are taken in passes of WNW x WPW blocks, WPW per wave (a
// wave of sixteen has 128 registers, and the 3M accumulators of two blocks
// take 48); each pass streams the operands again, from L2.  Workgroups are
// persistent over the bucket's tile list with k_tile's XCD-aware schedule.
constexpr int WNP = 128, WNW = 16, WTHR = WNW * 64, WK = 16, WPW = 2;
static_assert(WNP == G12_MAX_N, "wide tiles cover every N the library takes");
struct WideLds {
    static constexpr int PA = WK + 1, PB = WNP + 1;  // padded pitches (doubles)
    double ar[WNP * PA], ai[WNP * PA];               // A chunk [m][k]
    double br[WK * PB], bi[WK * PB];                 // B chunk [k][n]
};
// staged elements per thread and chunk: WNP x WK of A, WK x WNP of B
constexpr int WSLOTS = WNP * WK / WTHR;
static_assert(WSLOTS * WTHR == WNP * WK && WK == 16 && WNP == 128, "slot decomposition below");

template <bool FUSED>
__global__ __launch_bounds__(WTHR) void k_tile_wide(BTileArgs a) {
    __shared__ WideLds L;
    __shared__ double2 srcv[FUSED ? 1 : 3][WNP];
    constexpr int PA = WideLds::PA, PB = WideLds::PB;
    const StageShared ss = uni_stage(a.r);
    const TileSched sched(a.k.tiles, gridDim.x, blockIdx.x);
    const int wave = uni(threadIdx.x >> 6), lane = threadIdx.x & 63, tid = threadIdx.x;
    const int r16 = lane & 15, kq = lane >> 4;
    Cursor c;
    View v;
    for (int it = 0;; ++it) {
        const int64_t g = sched.tile(it, blockIdx.x, gridDim.x);
        if (g < 0) break;
        seek(a, ss, c, g);
        if (!c.on) continue;
        const TileRef tr = ref_of(c, g);
        const int N = tr.N;
        if (tr.traj != v.traj || tr.spin != v.spin) load_view<FUSED>(v, a, ss, tr.traj, tr.spin);
        const double2* h = uni_ptr(a.h_override ? a.h_override : a.b.ds[tr.traj].h);
        const double2* X = v.X;
        const TileGeom gn = tile_geom(FUSED ? 1 : a.tiling, N, tr.tile, false);
        const int t0 = gn.t0, t1 = gn.t1;
        const int nt = (N + 15) >> 4, TT = nt * nt;
        const int nc = (N + WK - 1) / WK;
        __syncthreads();  // the previous tile's readers of LDS are done
        if (a.sources && wave == 0) {  // source vectors, elements x = lane + 64 i
            for (int x = lane; x < N; x += 64) {
                SrcVec sv;
                src_load<FUSED>(v, tr, x, sv);
                if constexpr (FUSED) {
                    srcv[0][x] = cadd(cmul_i(sv.q), sv.f);
                } else {
                    srcv[0][x] = cadd(cscale(sv.g, sv.w0), cmul_i(sv.q));
                    srcv[1][x] = cadd(cscale(sv.g, sv.w1), cmul_i(sv.q));
                    srcv[2][x] = cconj(sv.f);
                }
            }
        }
        // chunks [0, nc): h.X; [nc, 2 nc): -X.h
        const int c0 = (a.hmode & 1) ? 0 : nc, c1 = (a.hmode & 2) ? 2 * nc : nc;
        // fused stages 1 and 2 read their input from Y and write the next
        // stage's input there: with more than one pass, a pass would overwrite
        // elements a later pass still streams, so each pass writes the next
        // input into R instead (R of an element is dead once this pass has
        // added it) and the tile copies R to Y after its last pass
        const bool park = FUSED && !ss.last && X == v.Y && TT > WNW * WPW;
        double2 ra[WSLOTS], rb[WSLOTS];
        auto fetch = [&](int ci) {
            const bool second = ci >= nc;
            const int k0 = (second ? ci - nc : ci) * WK;
#pragma unroll
            for (int i = 0; i < WSLOTS; ++i) {
                const int s = tid + WTHR * i;
                const int m = s >> 4, ka = k0 + (s & 15);   // A slot (m, k)
                const int kb = k0 + (s >> 7), n = s & 127;  // B slot (k, n)
                double2 va = make_double2(0.0, 0.0), vb = va;
                if (m < N && ka < N) {
                    if (second) {
                        const double2 x = gld(X, gn.nat(m, ka));
                        va = make_double2(-x.x, -x.y);
                    } else {
                        va = gld(h, (uint32_t)(m * N + ka));
                    }
                }
                if (kb < N && n < N) vb = second ? gld(h, (uint32_t)(kb * N + n)) : gld(X, gn.nat(kb, n));
                ra[i] = va;
                rb[i] = vb;
            }
        };
        for (int tb = 0; tb < TT; tb += WNW * WPW) {  // passes over the output blocks
            d4 p1[WPW], p2[WPW], p3[WPW];
#pragma unroll
            for (int u = 0; u < WPW; ++u) {
                p1[u] = d4{0.0, 0.0, 0.0, 0.0};
                p2[u] = p1[u];
                p3[u] = p1[u];
            }
            // R (the strided-pass result) enters the fused accumulators between
            // the two products, as in k_tile: Re = P1 - P2 through P1,
            // Im = P3 - P1 - P2 through P3
            auto mid = [&]() {
                if constexpr (FUSED) {
#pragma unroll
                    for (int u = 0; u < WPW; ++u) {
                        const int t = tb + wave + WNW * u;
                        if (t >= TT) continue;
                        const int n = (t % nt) * 16 + r16;
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int m = (t / nt) * 16 + 4 * j + kq;
                            if (m < N && n < N) {
                                const double2 rv = gld(v.R, gn.nat(m, n));
                                p1[u][j] += rv.x;
                                p3[u][j] += rv.x + rv.y;
                            }
                        }
                    }
                }
            };
            if (c0 < c1) fetch(c0);
            for (int ci = c0; ci < c1; ++ci) {
                if (ci == nc) mid();
                __syncthreads();  // the previous chunk's readers are done
#pragma unroll
                for (int i = 0; i < WSLOTS; ++i) {
                    const int s = tid + WTHR * i;
                    L.ar[(s >> 4) * PA + (s & 15)] = ra[i].x;
                    L.ai[(s >> 4) * PA + (s & 15)] = ra[i].y;
                    L.br[(s >> 7) * PB + (s & 127)] = rb[i].x;
                    L.bi[(s >> 7) * PB + (s & 127)] = rb[i].y;
                }
                __syncthreads();
                if (ci + 1 < c1) fetch(ci + 1);  // in flight during this chunk's MFMAs
#pragma unroll
                for (int j = 0; j < WK / 4; ++j) {
                    const int kk = kq * (WK / 4) + j;
#pragma unroll
                    for (int u = 0; u < WPW; ++u) {
                        const int t = tb + wave + WNW * u;
                        if (t >= TT) continue;
                        const int mm = (t / nt) * 16 + r16, nn = (t % nt) * 16 + r16;
                        const double xr = L.ar[mm * PA + kk], xi = L.ai[mm * PA + kk];
                        const double yr = L.br[kk * PB + nn], yi = L.bi[kk * PB + nn];
                        p1[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr, yr, p1[u], 0, 0, 0);
                        p2[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi, yi, p2[u], 0, 0, 0);
                        p3[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr + xi, yr + yi, p3[u], 0, 0, 0);
                    }
                }
            }
            if (c1 <= nc) mid();
            __syncthreads();  // srcv written by wave 0 is visible (no chunk may have run)
#pragma unroll
            for (int u = 0; u < WPW; ++u) {
                const int t = tb + wave + WNW * u;
                if (t >= TT) continue;
                const int n = (t % nt) * 16 + r16;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int m = (t / nt) * 16 + 4 * j + kq;
                    if (m >= N || n >= N) continue;
                    const double2 o = make_double2(p1[u][j] - p2[u][j], p3[u][j] - p1[u][j] - p2[u][j]);
                    const uint32_t off = gn.nat(m, n);
                    if constexpr (!FUSED) {
                        // strided-pass sources: g11 (a[m] - b[n]) - conj g(q1,p1) (c[n] - c[m])
                        double2 val = o;
                        if (a.sources) {
                            const uint32_t gi = (uint32_t)(m * N + n);
                            const double2 g11 = gld(v.G1, gi), gt = gld(v.G1t, gi);
                            val = cadd(val, csub(cmul(g11, csub(srcv[0][m], srcv[1][n])),
                                                 cmul(gt, csub(srcv[2][n], srcv[2][m]))));
                        }
                        if (a.accumulate) val = cadd(gld(v.R, off), val);
                        gst(v.R, off, val);
                    } else {
                        // fused update (k_tile): K = -i (R + O + g22 (e[m] - e[n]));
                        // A += w K and Y = S + c_next K, or S += dt (A + w K) / 6
                        double2 rr = o;
                        if (a.sources)
                            rr = cadd(rr, cmul(gld(v.G1, (uint32_t)(m * N + n)), csub(srcv[0][m], srcv[0][n])));
                        const double2 K = cmul_mi(rr);
                        const double2 wK = cscale(K, ss.wk);
                        const double2 sv = gld(v.S, off);
                        double2 next;
                        if (!ss.last) {
                            gst(v.A, off, ss.first ? wK : cadd(gld(v.A, off), wK));
                            next = cadd(sv, cscale(K, ss.cn));
                            gst(park ? v.R : v.Y, off, next);
                        } else {
                            next = cadd(sv, cdiv6(cscale(cadd(gld(v.A, off), wK), ss.dt)));
                            gst(v.S, off, next);
                        }
                        if (m == n) gst(v.D, (uint32_t)((t0 * N + t1) * N + m), next);
                    }
                }
            }
        }
        if (park) {
            __syncthreads();  // every pass has streamed X (= Y)
            // each thread moves the elements it wrote itself
            for (int tb = 0; tb < TT; tb += WNW * WPW)
#pragma unroll
                for (int u = 0; u < WPW; ++u) {
                    const int t = tb + wave + WNW * u;
                    if (t >= TT) continue;
                    const int n = (t % nt) * 16 + r16;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int m = (t / nt) * 16 + 4 * j + kq;
                        if (m < N && n < N) gst(v.Y, gn.nat(m, n), gld(v.R, gn.nat(m, n)));
                    }
                }
        }
    }
}

// Start of a run: partial traces of the state (the stage-0 input)
// D[X][Y][Z] = S[X][Z][Y][Z] for both spins and, in mixed-precision runs, the
// complex64 copy Sc of S that the tile passes read as the stage-0 input;
// within a run each fused pass writes both for the next stage input.  One
// block per slice (p1, q1) = (X, Y), blocks in the bucket's tile order (spin 0
// tiles).  a.sym: S (and Sc) in the Hermitian half layout, where
// S[X][Z][Y][Z] is element (Z, Z) of block (X, Y), or for X > Y the conjugate
// of that of block (Y, X).
__global__ __launch_bounds__(64) void k_d_from_state(BTileArgs a) {
    const int64_t b = blockIdx.x;
    int lo = 0, hi = a.k.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.k.prefix[mid] / 2 <= b) lo = mid; else hi = mid - 1;
    }
    const int traj = a.k.list[lo];
    const DevState& d = a.b.ds[traj];
    const int N = a.b.tp[traj].N;
    const int tile = (int)(b - a.k.prefix[lo] / 2);
    if (tile >= N * N) return;
    const int X = tile / N, Y = tile % N;
    const int64_t N2 = (int64_t)N * N;
    if (a.sym) {
        const int lo2 = X < Y ? X : Y, hi2 = X < Y ? Y : X;
        const int64_t base = (int64_t)hblock(lo2, hi2) * N2;
        for (int Z = threadIdx.x; Z < N; Z += blockDim.x)
            for (int s = 0; s < 2; ++s) {
                const double2 v = gload(d.S2[s], base + (int64_t)Z * N + Z);
                gstore(d.D[s], ((int64_t)X * N + Y) * N + Z, X <= Y ? v : cconj(v));
            }
        if (a.compact && X <= Y)
            for (int s = 0; s < 2; ++s)
                for (int64_t e = threadIdx.x; e < N2; e += blockDim.x)
                    cstore(d.Sc[s], base + e, gload(d.S2[s], base + e), true);
        return;
    }
    for (int Z = threadIdx.x; Z < N; Z += blockDim.x) {
        const int64_t src = (int64_t)X * N2 * N + (int64_t)Z * N2 + (int64_t)Y * N + Z;
        const int64_t dst = ((int64_t)X * N + Y) * N + Z;
        for (int s = 0; s < 2; ++s) gstore(d.D[s], dst, gload(d.S2[s], src));
    }
    if (a.compact) {  // the contiguous tile (p1, q1) = (X, Y): elements (p2, q2), blocked layout
        const TileGeom gc = tile_geom(1, N, tile, BLK_Y);
        for (int s = 0; s < 2; ++s)
            for (int p2 = 0; p2 < N; ++p2)
                for (int q2 = threadIdx.x; q2 < N; q2 += blockDim.x) {
                    const int64_t e = (int64_t)X * N2 * N + (int64_t)p2 * N2 + (int64_t)Y * N + q2;
                    cst(d.Sc[s], gc.at<BLK_Y>(p2, q2), gload(d.S2[s], e), true);
                }
    }
}

// ---------------------------------------------------------------------------
// Pair-symmetric path, second pass.  A physical G2 tensor is symmetric under
// the exchange of both pairs, T[p1][p2][q1][q2] = T[p2][p1][q2][q1], and
// Hermitian as an N^2 x N^2 matrix, with G1 anti-Hermitian (every state
// evolved from G2 = 0 is).  Its right-hand side is then pair-symmetric too:
//   [H2, M] + Psi + Pi - Pi^+ = X + P(X),   P(X)[p1][p2][q1][q2] = X[p2][p1][q2][q1],
// with X what the SYM tile pass writes (k_tile): the p1, q1 terms of [H2, M]
// are the exchanged p2, q2 terms, the Psi term i g11 (Qp[p2][q2][p1] -
// Qp[p2][q2][q1]) the exchanged g22 term, Pi^+ the exchanged Pi (F is
// Hermitian in its first two indices when G2 is Hermitian), and the Psi term
// g11 g22 (W[p1][p2] - W[q1][q2]) is its own exchange, half of it in X.
//
// Every buffer of the pass is in the Hermitian half layout (hblock).  The
// stored elements split into the direct and the mirrored family, each a
// blocked transpose of units (sym_units.h); units are packed into blocks of
// 1024 slots on the host (su_table).  Each thread takes one element of each
// side of its unit; both sides go through LDS.  Then the RK update of every
// element with K = -i C, the partial traces D of the next stage input, and the
// Hermitian images in the diagonal blocks.
template <bool C>
__device__ __forceinline__ void sym_update_body(const BatchDev& b, const StageRef& r) {
    typedef typename std::conditional<C, fv2, dv2>::type LT;
    __shared__ LT la[SU_LA];
    __shared__ LT lb[SU_LB];
    __shared__ uint64_t ud[SU_MAXU];
    const StageShared ss = uni_stage(r);
    const int64_t blk = blockIdx.x;
    int lo = 0, hi = b.ncorr - 1;
    while (lo < hi) {  // last correlated trajectory whose blocks start at or before this one
        const int mid = (lo + hi + 1) >> 1;
        if (b.sprefix[mid] <= blk) lo = mid; else hi = mid - 1;
    }
    const int traj = uni(b.tlist[lo]);
    const DevState& d = b.ds[traj];
    if (!active(d, b.tp[traj], ss)) return;
    const int N = uni(b.tp[traj].N);
    const int nblk = uni(d.su_nblk);
    int64_t bb = blk - uni64(b.sprefix[lo]);
    const int spin = bb >= nblk ? 1 : 0;  // blocks never mix spins
    bb -= spin * nblk;
    const uint32_t* first = uni_ptr(d.su_first);
    const uint32_t f0 = (uint32_t)uni((int)first[bb]), f1 = (uint32_t)uni((int)first[bb + 1]);
    const int cnt = (int)(f1 - f0);
    if ((int)threadIdx.x < cnt) ud[threadIdx.x] = uni_ptr(d.su_unit)[f0 + threadIdx.x];
    __syncthreads();
    // this thread's unit: the last one whose slots start at or before it
    const int t = threadIdx.x;
    const SuUnit u = su_decode(ud[su_find(ud, cnt, t)], N);
    const int e = t - u.so;
    const bool live = e < u.R * u.L;
    const double2* X = uni_ptr((const double2*)d.R2[spin]);
    double2* S = uni_ptr(d.S2[spin]);
    double2* A = uni_ptr(d.A2[spin]);
    double2* Y = uni_ptr(d.Y2[spin]);
    double2* Sc = uni_ptr(d.Sc[spin]);  // mixed runs: complex64 copy of S
    double2* D = uni_ptr(d.D[spin]);
    const bool mir = u.type != 0;
    // the rest in two forms, with and without the non-temporal accesses (a
    // run-time choice per access would let the compiler merge the two and
    // drop the non-temporal hint); the choice is uniform over the block
    auto body = [&](auto ntc) {
        constexpr bool nt = decltype(ntc)::value;
        const SuElem q = su_elem(u, live ? e : 0, N);
        // RK protocol of this pass (its stages see S and A only at element level).
        // Mixed runs: stages 0 and 1 store their K (complex64, the precision of the
        // FP32 products it comes from) in the A buffer, as two complex64 arrays
        // K1, K2, and write Y = Sc + c' K (Sc: the complex64 copy of S; the stage
        // inputs are complex64 anyway); stage 2 writes Y = S + c' K and folds
        // S += dt (K1 + 2 K2 + 2 K3) / 6 in FP64; stage 3 S += dt K4 / 6, and Sc.
        // FP64 runs keep no accumulator: stage 0 writes the stage-1 input Y1 to Y,
        // stage 1 the stage-2 input Y2 to the A buffer, and stage 2 recovers
        // K1 = (Y1 - S) / c and K2 = (Y2 - S) / c (c = dt/2 for both; exact to the
        // rounding of Y, a relative 1e-16 of S), folds S += dt (K1 + 2 K2 + 2 K3) / 6
        // and writes the stage-3 input over Y1.
        const bool s_c64 = C && !ss.fold && !ss.last;
        const bool fold64 = !C && ss.fold;  // FP64, stage 2
        const uint32_t k2o = hblock(0, N) * (uint32_t)(N * N);  // mixed: K2 after K1 in the A buffer (N (N + 1) / 2 blocks)
        struct Item {
            uint32_t o;
            bool ok;
            LT xv;       // this element of X
            double2 sv;  // S (or Sc)
            LT av, a2;   // stage 2, as stored: K1, K2 (mixed) or Y1, Y2 (FP64)
        };
        auto gather = [&](uint32_t o, Item& w, bool with_state) {
            w.ok = live;
            w.o = o;
            w.sv = make_double2(0.0, 0.0);
            w.av = w.a2 = LT{0, 0};
            if (w.ok) {
                if constexpr (C && nt)
                    w.xv = __builtin_nontemporal_load((const GLOBAL fv2*)((const GLOBAL char*)X + (o << 3)));
                else if constexpr (C)
                    w.xv = *(const GLOBAL fv2*)((const GLOBAL char*)X + (o << 3));
                else
                    w.xv = __builtin_nontemporal_load((const GLOBAL dv2*)((const GLOBAL char*)X + (o << 4)));
                if (with_state) {
                    w.sv = s_c64 ? cld(Sc, o, true) : gld(S, o);
                    if constexpr (C) {
                        if (ss.fold) {
                            w.av = *(const GLOBAL fv2*)((const GLOBAL char*)A + (o << 3));          // K1
                            w.a2 = *(const GLOBAL fv2*)((const GLOBAL char*)A + ((k2o + o) << 3));  // K2
                        }
                    } else if (fold64) {
                        w.av = *(const GLOBAL dv2*)((const GLOBAL char*)Y + (o << 4));  // Y1
                        w.a2 = *(const GLOBAL dv2*)((const GLOBAL char*)A + (o << 4));  // Y2
                    }
                }
            }
        };
        // all loads of the block (X and the RK state of both sides) in flight at once
        Item w1, w2;
        gather(q.o1, w1, true);
        gather(q.o2, w2, q.w2);  // x = y: side 2 is read for X only
        LT* lau = la + u.la;
        LT* lbu = lb + u.lb;
        if (live) {
            lau[q.i1 * (u.L + 1) + q.j1] = w1.xv;
            lbu[q.j2 * (u.R + 1) + q.i2] = w2.xv;
        }
        __syncthreads();
        if (!live) return;
        auto tod = [](LT v) { return make_double2((double)v.x, (double)v.y); };
        // a complex64 or complex128 store (non-temporal) and, when there is one,
        // its Hermitian image
        auto put = [&](double2* p, bool c64, uint32_t o, uint32_t im, double2 v) {
            if constexpr (nt)
                cst_nt(p, o, v, c64);
            else
                cst(p, o, v, c64);
            if (im != SU_NONE) cst(p, im, cconj(v), c64);
        };
        auto update = [&](const Item& w, uint32_t im, int bp, int bq, int z, double2 c) {
            const double2 K = cmul_mi(c);
            const double2 wK = cscale(K, ss.wk);
            const double2 v1 = tod(w.av), v2 = tod(w.a2);
            double2 av = v1;  // at stage 2: K1 + 2 K2
            if (fold64) {
                const double2 k1 = make_double2((v1.x - w.sv.x) / ss.c, (v1.y - w.sv.y) / ss.c);
                const double2 k2 = make_double2((v2.x - w.sv.x) / ss.c, (v2.y - w.sv.y) / ss.c);
                av = cadd(k1, cscale(k2, 2.0));
            } else if (C && ss.fold) {
                av = cadd(v1, cscale(v2, 2.0));
            }
            double2 next;
            if (ss.last) {
                next = cadd(w.sv, cdiv6(cscale(wK, ss.dt)));
                put(S, false, w.o, im, next);
                if constexpr (C) put(Sc, true, w.o, im, next);  // next step's stage-0 input
            } else {
                next = cadd(w.sv, cscale(K, ss.cn));
                if (!C && !ss.first && !ss.fold)
                    put(A, false, w.o, im, next);  // FP64 stage 1: the stage-2 input into the A buffer
                else
                    put(Y, C, w.o, im, next);
                if (ss.fold)
                    put(S, false, w.o, im, cadd(w.sv, cdiv6(cscale(cadd(av, wK), ss.dt))));
                else if constexpr (C)
                    put(A, true, (ss.first ? 0u : k2o) + w.o, im == SU_NONE ? SU_NONE : (ss.first ? 0u : k2o) + im, K);  // K1 or K2
            }
            if (z >= 0) {  // D[bp][bq][z] = T[bp][z][bq][z], and its Hermitian mirror
                gst(D, ((uint32_t)bp * N + bq) * N + z, next);
                if (bp != bq) gst(D, ((uint32_t)bq * N + bp) * N + z, cconj(next));
            }
        };
        // C = X + exchanged X: the partner from LDS, as it is (direct) or -conj (mirrored)
        auto pair = [&](LT v) {
            const double2 p = tod(v);
            return mir ? make_double2(-p.x, p.y) : p;
        };
        update(w1, q.im1, q.bp1, q.bq1, q.z1, cadd(tod(w1.xv), pair(lbu[q.j1 * (u.R + 1) + q.i1])));
        if (q.w2) update(w2, q.im2, q.bp2, q.bq2, q.z2, cadd(tod(w2.xv), pair(lau[q.i2 * (u.L + 1) + q.j2])));
    };
    if (use_nt(C, N))
        body(std::true_type{});
    else
        body(std::false_type{});
}

// The pass as a kernel.  The FP64 form fits 64 VGPRs, two 1024-thread blocks
// per CU.  The complex64 form needs 68; held to 64 (8 waves per SIMD, two
// small spills) it ran the N = 64 mixed update 10% faster, but the 256-draw
// ensemble (N = 6..60, most blocks on the cached-store form) 2.6% slower, so
// both are built and a launch takes the bounded one when every trajectory's
// rows fill whole lines (k_sym_update_w8).
template <bool C>
__global__ __launch_bounds__(SU_THREADS) void k_sym_update(BatchDev b, StageRef r) {
    sym_update_body<C>(b, r);
}
__global__ __launch_bounds__(SU_THREADS) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_sym_update_w8(BatchDev b,
                                                                                                     StageRef r) {
    sym_update_body<true>(b, r);
}

// flag |= 1 unless T is pair-symmetric (and, with herm, Hermitian) within tol,
// elementwise
__global__ void k_pair_check(const double2* T, int N, double tol, int herm, int* flag) {
    const int64_t n = (int64_t)N * N, n4 = n * n;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n4) return;
    const int q2 = (int)(e % N), q1 = (int)((e / N) % N), p2 = (int)((e / n) % N), p1 = (int)(e / (n * N));
    const double2 v = gload(T, e);
    const double2 vp = gload(T, (((int64_t)p2 * N + p1) * N + q2) * N + q1);
    const double2 vh = gload(T, (((int64_t)q1 * N + q2) * N + p1) * N + p2);
    const double dp = fmax(fabs(v.x - vp.x), fabs(v.y - vp.y));
    const double dh = herm ? fmax(fabs(v.x - vh.x), fabs(v.y + vh.y)) : 0.0;
    if (!(dp <= tol && dh <= tol)) atomicOr(flag, 1);
}

int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return cus;
}

template <int NP, int NW, int PREC, bool FULL>
void launch_tile_cfg(const BTileArgs& a, hipStream_t s) {
    // the engine's tile images plus the source vectors (3 x NP complex at most)
    constexpr int lds = (int)sizeof(typename Engine<NP, NW, PREC>::Lds) + 3 * NP * (int)sizeof(double2);
    int per_cu = (160 * 1024) / lds;
    const int by_waves = 32 / NW;  // at most 32 waves per CU
    if (per_cu > by_waves) per_cu = by_waves;
    if (a.k.wg_per_cu > 0 && per_cu > a.k.wg_per_cu) per_cu = a.k.wg_per_cu;
    if (per_cu < 1) per_cu = 1;
    int64_t grid = (int64_t)device_cus() * per_cu;
    if (grid > a.k.tiles) grid = a.k.tiles;
    if (grid >= 8) grid &= ~7;
    if (grid < 1) return;
    // the full-tile form serves the fused pass in FP32-3M only: in FP64 it
    // spills registers (9% more HBM reads) and ran no faster, in BF16x3 it ran
    // 20% slower
    constexpr bool FULL_FUSED = FULL && PREC == G12_PREC_FP32_3M;
    if (a.fused)
        hipLaunchKernelGGL((k_tile<NP, NW, PREC, true, FULL_FUSED>), dim3((unsigned)grid),
                           dim3(TileCfg<NP, NW>::NTHR), 0, s, a);
    else if (a.sym)  // BF16x3's full-tile form of this pass spills registers
        hipLaunchKernelGGL((k_tile<NP, NW, PREC, false, FULL && PREC != G12_PREC_BF16X3, true>), dim3((unsigned)grid),
                           dim3(TileCfg<NP, NW>::NTHR), 0, s, a);
    else
        hipLaunchKernelGGL((k_tile<NP, NW, PREC, false, FULL>), dim3((unsigned)grid), dim3(TileCfg<NP, NW>::NTHR), 0,
                           s, a);
}

// Wave counts compiled for a padded tile size NP: W0, 2 W0 and 4 W0 with
// W0 = 1, 2, 4 for NP = 16, 32, >= 48 (a tile has (NP/16)^2 output blocks of
// 16x16: one at NP = 16, sixteen at NP = 64).
constexpr int base_waves(int NP) { return NP == 16 ? 1 : NP == 32 ? 2 : 4; }

template <int NP, int PREC>
void launch_tile_np(const BTileArgs& a, hipStream_t s) {
    constexpr int W0 = base_waves(NP);
    if (a.k.waves >= 4 * W0) {
        // buckets whose trajectories all have N = NP take the full-tile
        // kernels (for the fused pass where they measured faster)
        if (a.k.full)
            launch_tile_cfg<NP, 4 * W0, PREC, true>(a, s);
        else
            launch_tile_cfg<NP, 4 * W0, PREC, false>(a, s);
    } else if (a.k.waves >= 2 * W0) {
        launch_tile_cfg<NP, 2 * W0, PREC, false>(a, s);
    } else {
        launch_tile_cfg<NP, W0, PREC, false>(a, s);
    }
}

// Wide tiles (NP = 80 ... 128)
void launch_tile_wide(const BTileArgs& a, hipStream_t s) {
    constexpr int lds = (int)sizeof(WideLds) + 3 * WNP * (int)sizeof(double2);
    int per_cu = (160 * 1024) / lds;
    if (per_cu > 32 / WNW) per_cu = 32 / WNW;
    if (a.k.wg_per_cu > 0 && per_cu > a.k.wg_per_cu) per_cu = a.k.wg_per_cu;
    int64_t grid = (int64_t)device_cus() * per_cu;
    if (grid > a.k.tiles) grid = a.k.tiles;
    if (grid >= 8) grid &= ~7;
    if (grid < 1) return;
    if (a.fused)
        hipLaunchKernelGGL(k_tile_wide<true>, dim3((unsigned)grid), dim3(WTHR), 0, s, a);
    else
        hipLaunchKernelGGL(k_tile_wide<false>, dim3((unsigned)grid), dim3(WTHR), 0, s, a);
}

template <int PREC>
void launch_tile_prec(const BTileArgs& a, hipStream_t s) {
    switch (a.k.NP) {
        case 16: launch_tile_np<16, PREC>(a, s); break;
        case 32: launch_tile_np<32, PREC>(a, s); break;
        case 48: launch_tile_np<48, PREC>(a, s); break;
        case 64: launch_tile_np<64, PREC>(a, s); break;
        default:
            // N > 64: FP64 only (the host rejects mixed precision there)
            if constexpr (PREC == G12_PREC_FP64) launch_tile_wide(a, s);
            break;
    }
}

void launch_tile(const BTileArgs& a, int precision, hipStream_t s) {
    if (a.k.tiles <= 0) return;
    if (precision == G12_PREC_FP32_3M)
        launch_tile_prec<G12_PREC_FP32_3M>(a, s);
    else if (precision == G12_PREC_BF16X3)
        launch_tile_prec<G12_PREC_BF16X3>(a, s);
    else
        launch_tile_prec<G12_PREC_FP64>(a, s);
}

// ---------------------------------------------------------------------------
// Layout permutations between the reference pair order and the natural one.
__device__ __forceinline__ int64_t ref_pair(int p1, int p2, int Ns, int Nb) {
    const int i = p1 % Ns, a = p1 / Ns, j = p2 % Ns, b = p2 / Ns;
    return (int64_t)i + (int64_t)j * Ns + ((int64_t)a + (int64_t)b * Nb) * Ns * Ns;
}

__global__ void k_ref_to_nat(const double2* in, double2* out, int Ns, int Nb) {
    const int N = Ns * Nb;
    const int64_t N2 = (int64_t)N * N;
    const int64_t A = blockIdx.x;  // natural row p1*N + p2
    const int p1 = (int)(A / N), p2 = (int)(A % N);
    const int64_t Ar = ref_pair(p1, p2, Ns, Nb);
    for (int B = threadIdx.x; B < N2; B += blockDim.x) {
        const int q1 = B / N, q2 = B % N;
        out[A * N2 + B] = in[Ar * N2 + ref_pair(q1, q2, Ns, Nb)];
    }
}

__global__ void k_nat_to_ref(const double2* in, double2* out, int Ns, int Nb) {
    const int N = Ns * Nb;
    const int64_t N2 = (int64_t)N * N;
    const int64_t A = blockIdx.x;
    const int p1 = (int)(A / N), p2 = (int)(A % N);
    const int64_t Ar = ref_pair(p1, p2, Ns, Nb);
    for (int B = threadIdx.x; B < N2; B += blockDim.x) {
        const int q1 = B / N, q2 = B % N;
        out[Ar * N2 + ref_pair(q1, q2, Ns, Nb)] = in[A * N2 + B];
    }
}

// ---------------------------------------------------------------------------
// Routine-level kernels (reference-routine entry points).
__global__ void k_comm1(const double2* G, const double2* h, double2* out, int N) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * N) return;
    const int X = e / N, Y = e % N;
    double2 hg = make_double2(0.0, 0.0), gh = make_double2(0.0, 0.0);
    for (int z = 0; z < N; ++z) {
        hg = cadd(hg, cmul(h[(int64_t)X * N + z], G[(int64_t)z * N + Y]));
        gh = cadd(gh, cmul(G[(int64_t)X * N + z], h[(int64_t)z * N + Y]));
    }
    out[e] = csub(hg, gh);
}

// Pair-diagonal elements M[(X,Y);(X,Y)] = T[X][Y][X][Y] of a G2 tensor in the
// natural layout, or (half) element (Y, Y) of block (X, X) of the Hermitian
// half layout (the correlation energy's only input from G2).
__global__ void k_pair_diag(const double2* T, double2* out, int N, int half) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t N2 = (int64_t)N * N;
    if (e >= N2) return;
    const int64_t X = e / N, Y = e % N;
    out[e] = half ? T[(int64_t)hblock((int)X, (int)X) * N2 + Y * N + Y] : T[((X * N + Y) * N + X) * N + Y];
}

// natural -> Hermitian half layout: one block per tile (p1, q1), p1 <= q1;
// a diagonal tile takes its upper triangle and the conjugate transpose of it
__global__ void k_nat_to_half(const double2* nat, double2* half, int N) {
    const int q1 = blockIdx.x / N, p1 = blockIdx.x % N;
    if (p1 > q1) return;
    const int64_t N2 = (int64_t)N * N;
    const int64_t out0 = (int64_t)hblock(p1, q1) * N2;
    for (int64_t e = threadIdx.x; e < N2; e += blockDim.x) {
        const int p2 = (int)(e / N), q2 = (int)(e % N);
        const bool low = p1 == q1 && p2 > q2;  // take (q2, p2) conjugated
        const int r = low ? q2 : p2, c = low ? p2 : q2;
        const double2 v = nat[(((int64_t)p1 * N + r) * N + q1) * N + c];
        half[out0 + e] = low ? cconj(v) : v;
    }
}

// Hermitian half -> natural layout: one block per natural slice (p1, p2)
__global__ void k_half_to_nat(const double2* half, double2* nat, int N) {
    const int p1 = blockIdx.x / N, p2 = blockIdx.x % N;
    const int64_t N2 = (int64_t)N * N;
    for (int64_t e = threadIdx.x; e < N2; e += blockDim.x) {
        const int q1 = (int)(e / N), q2 = (int)(e % N);
        const double2 v = p1 <= q1 ? half[(int64_t)hblock(p1, q1) * N2 + (int64_t)p2 * N + q2]
                                   : cconj(half[(int64_t)hblock(q1, p1) * N2 + (int64_t)q2 * N + p2]);
        nat[(int64_t)blockIdx.x * N2 + e] = v;
    }
}

__global__ void k_g1_step(const double2* comm, const double2* I, double2* K, double2* RK, double w, int N) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * N) return;
    const int X = e / N, Y = e % N;
    const double2 k = cmul_mi(cadd(cadd(comm[e], I[e]), cconj(I[(int64_t)Y * N + X])));
    K[e] = k;
    RK[e] = cadd(RK[e], cscale(k, w));
}

// G2_step sum (G1_G2_t_step.h:55-86) of the natural layout, whose
// particle-swapped transposed partner of (p1,p2,q1,q2) is (q2,q1,p2,p1):
// out = comm + psi + pi - conj(pi[partner])
__global__ void k_g2_sum(const double2* comm, const double2* psi, const double2* pi, double2* out, int N) {
    const int64_t N2 = (int64_t)N * N;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N2 * N2) return;
    const int p1 = (int)(e / (N2 * N)), p2 = (int)((e / N2) % N), q1 = (int)((e / N) % N), q2 = (int)(e % N);
    const int64_t t = (((int64_t)q2 * N + q1) * N + p2) * N + p1;
    out[e] = csub(cadd(cadd(comm[e], psi[e]), pi[e]), cconj(pi[t]));
}

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------------------
// launchers

void launch_b_stage_setup(const BatchDev& b, const StageRef& r, int precision, hipStream_t s) {
    const int gA = (b.maxN * b.maxN + SETUP_THREADS - 1) / SETUP_THREADS;
    const int64_t blocks = (int64_t)gA * b.ntraj + 2 * b.tblocks;
    hipLaunchKernelGGL(k_stage_setup, dim3((unsigned)blocks), dim3(SETUP_THREADS), 0, s, b, r,
                       precision != G12_PREC_FP64 ? 1 : 0, gA);
}

void launch_b_g1_rhs(const BatchDev& b, const StageRef& r, hipStream_t s) {
    hipLaunchKernelGGL(k_g1_rhs, dim3(b.maxN, b.ntraj), dim3(256), 0, s, b, r);
}

void launch_b_record(const BatchDev& b, const StageRef& r, hipStream_t s) {
    const unsigned nb = (unsigned)((b.maxN * b.maxN + 255) / 256);
    hipLaunchKernelGGL(k_record, dim3(nb, b.ntraj), dim3(256), 0, s, b, r);
}

void launch_b_advance_counter(int* counter, hipStream_t s) {
    hipLaunchKernelGGL(k_advance_counter, dim3(1), dim3(64), 0, s, counter);
}

int device_cu_count() { return device_cus(); }

int default_waves(int NP) { return NP >= 48 ? 16 : 4; }

int tile_wave_options(int NP, int* opts) {
    if (NP > 64) {  // wide tiles: one wave count
        opts[0] = WNW;
        return 1;
    }
    const int w0 = base_waves(NP);
    opts[0] = w0;
    opts[1] = 2 * w0;
    opts[2] = 4 * w0;
    return 3;
}

void launch_b_d_from_state(const BatchDev& b, const Bucket& k, int precision, hipStream_t s, int half) {
    if (k.tiles <= 0) return;
    BTileArgs a{};
    a.b = b;
    a.k = k;
    a.compact = precision != G12_PREC_FP64;
    a.sym = half;
    hipLaunchKernelGGL(k_d_from_state, dim3((unsigned)(k.tiles / 2)), dim3(64), 0, s, a);
}

void launch_b_sym_update(const BatchDev& b, const StageRef& r, int precision, hipStream_t s) {
    if (b.sunits <= 0) return;
    if (precision != G12_PREC_FP64 && b.sym_w8)
        hipLaunchKernelGGL(k_sym_update_w8, dim3((unsigned)b.sunits), dim3(SU_THREADS), 0, s, b, r);
    else if (precision != G12_PREC_FP64)
        hipLaunchKernelGGL(k_sym_update<true>, dim3((unsigned)b.sunits), dim3(SU_THREADS), 0, s, b, r);
    else
        hipLaunchKernelGGL(k_sym_update<false>, dim3((unsigned)b.sunits), dim3(SU_THREADS), 0, s, b, r);
}

void launch_pair_check(const double2* T, int N, double tol, int herm, int* flag, hipStream_t s) {
    const int64_t n4 = (int64_t)N * N * N * N;
    hipLaunchKernelGGL(k_pair_check, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, s, T, N, tol, herm, flag);
}

void launch_b_tile(const BatchDev& b, const Bucket& k, const StageRef& r, int fused, int precision, hipStream_t s,
                   int sym) {
    BTileArgs a{};
    a.b = b;
    a.k = k;
    if (fused) {
        a.k.waves = k.waves_fused;
        a.k.wg_per_cu = k.wg_per_cu_fused;
    }
    a.r = r;
    a.sym = (!fused && sym) ? 1 : 0;
    if (a.sym) {  // the half tile list (t0 <= t1)
        a.k.prefix = k.hprefix;
        a.k.tiles = k.htiles;
    }
    a.tiling = (fused || a.sym) ? 1 : 0;
    a.fused = fused;
    a.hmode = 3;
    a.accumulate = 0;
    a.nspin = 2;
    a.use_y = 1;
    a.compact = precision != G12_PREC_FP64;
    a.sources = 1;
    a.h_override = nullptr;
    launch_tile(a, precision, s);
}

void launch_b_tile_args(const BatchDev& b, const Bucket& k, const StageRef& r, int fused, int hmode, int sources,
                        int accumulate, hipStream_t s) {
    BTileArgs a{};
    a.b = b;
    a.k = k;
    a.r = r;
    a.tiling = fused ? 1 : 0;
    a.fused = fused;
    a.hmode = hmode;
    a.accumulate = accumulate;
    a.nspin = 2;
    a.use_y = 1;
    a.compact = 0;
    a.sources = sources;
    a.h_override = nullptr;
    launch_tile(a, G12_PREC_FP64, s);
}

void launch_ref_to_nat(const double2* in, double2* out, int Ns, int Nb, hipStream_t s) {
    const int N = Ns * Nb;
    hipLaunchKernelGGL(k_ref_to_nat, dim3(N * N), dim3(256), 0, s, in, out, Ns, Nb);
}

void launch_nat_to_ref(const double2* in, double2* out, int Ns, int Nb, hipStream_t s) {
    const int N = Ns * Nb;
    hipLaunchKernelGGL(k_nat_to_ref, dim3(N * N), dim3(256), 0, s, in, out, Ns, Nb);
}

// Launch descriptors of a one-trajectory batch in a device area of
// COMM2_DESC_BYTES: DevState at 0, TrajParams at 512, the bucket list at 3072,
// the tile prefix at 3328.  Written by a kernel (kernel arguments by value),
// so the whole enqueue needs no host copy and can be captured in a graph.
constexpr int COMM2_DS = 0, COMM2_TP = 512, COMM2_LIST = 3072, COMM2_PREFIX = 3328;
static_assert(sizeof(DevState) <= COMM2_TP - COMM2_DS, "DevState fits its slot");
static_assert(COMM2_TP + sizeof(TrajParams) <= COMM2_LIST, "TrajParams fits its slot");
static_assert(COMM2_PREFIX + 2 * sizeof(int64_t) <= COMM2_DESC_BYTES, "descriptors fit");

static __global__ void k_comm2_desc(char* desc, DevState d, TrajParams tp, int64_t tiles) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *reinterpret_cast<DevState*>(desc + COMM2_DS) = d;
    *reinterpret_cast<TrajParams*>(desc + COMM2_TP) = tp;
    *reinterpret_cast<int*>(desc + COMM2_LIST) = 0;
    int64_t* pre = reinterpret_cast<int64_t*>(desc + COMM2_PREFIX);
    pre[0] = 0;
    pre[1] = tiles;
}

int enqueue_comm2(const double2* T, const double2* hx, const double2* hy, double2* out, int N, int precision,
                  void* desc_area, hipStream_t s) {
    // a batch of one trajectory whose spin-0 state is T and strided-pass
    // output is `out`; four accumulating launches (strided, contiguous) x
    // (left hx, right hy)
    DevState d{};
    d.N = N;
    d.S2[0] = const_cast<double2*>(T);
    d.R2[0] = out;
    TrajParams tp{};
    tp.N = N;
    tp.corr = 1;
    char* desc = static_cast<char*>(desc_area);
    hipLaunchKernelGGL(k_comm2_desc, dim3(1), dim3(64), 0, s, desc, d, tp, (int64_t)N * N);
    BTileArgs a{};
    a.b.ds = reinterpret_cast<const DevState*>(desc + COMM2_DS);
    a.b.tp = reinterpret_cast<const TrajParams*>(desc + COMM2_TP);
    a.b.ntraj = 1;
    a.b.maxN = N;
    a.k.NP = (N + 15) / 16 * 16;
    a.k.list = reinterpret_cast<const int*>(desc + COMM2_LIST);
    a.k.prefix = reinterpret_cast<const int64_t*>(desc + COMM2_PREFIX);
    a.k.n = 1;
    a.k.tiles = (int64_t)N * N;
    a.k.waves = default_waves(a.k.NP);
    a.k.waves_fused = a.k.waves;
    a.r.val.first = 1;
    a.nspin = 1;
    a.use_y = 0;
    a.compact = 0;
    a.fused = 0;
    int first = 1;
    for (int tiling = 0; tiling < 2; ++tiling)
        for (int side = 1; side <= 2; ++side) {
            a.tiling = tiling;
            a.hmode = side;
            a.h_override = side == 1 ? hx : hy;
            a.accumulate = first ? 0 : 1;
            first = 0;
            launch_tile(a, precision, s);
        }
    return hipGetLastError() == hipSuccess ? 0 : G12_E_HIP;
}

int launch_comm2(const double2* T, const double2* hx, const double2* hy, double2* out, int N, int precision,
                 hipStream_t s) {
    void* desc = nullptr;
    if (hipMalloc(&desc, COMM2_DESC_BYTES) != hipSuccess) return G12_E_NOMEM;
    const int launched = enqueue_comm2(T, hx, hy, out, N, precision, desc, s);
    const hipError_t e = hipStreamSynchronize(s);  // the tiles read the descriptors until here
    (void)hipFree(desc);
    return (launched == 0 && e == hipSuccess) ? 0 : G12_E_HIP;
}

void launch_comm1(const double2* G, const double2* h, double2* out, int N, hipStream_t s) {
    hipLaunchKernelGGL(k_comm1, dim3(blocks_for((int64_t)N * N, 256)), dim3(256), 0, s, G, h, out, N);
}

void launch_pair_diag(const double2* T, double2* out, int N, hipStream_t s, int half) {
    hipLaunchKernelGGL(k_pair_diag, dim3(blocks_for((int64_t)N * N, 256)), dim3(256), 0, s, T, out, N, half);
}

void launch_nat_to_half(const double2* nat, double2* half, int N, hipStream_t s) {
    hipLaunchKernelGGL(k_nat_to_half, dim3((unsigned)(N * N)), dim3(256), 0, s, nat, half, N);
}

void launch_half_to_nat(const double2* half, double2* nat, int N, hipStream_t s) {
    hipLaunchKernelGGL(k_half_to_nat, dim3((unsigned)(N * N)), dim3(256), 0, s, half, nat, N);
}

void launch_g1_step(const double2* comm, const double2* I, double2* K, double2* RK, double w, int N,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_g1_step, dim3(blocks_for((int64_t)N * N, 256)), dim3(256), 0, s, comm, I, K, RK, w, N);
}

void launch_g2_sum(const double2* comm, const double2* psi, const double2* pi, double2* out, int N, hipStream_t s) {
    const int64_t n4 = (int64_t)N * N * N * N;
    hipLaunchKernelGGL(k_g2_sum, dim3(blocks_for(n4, 256)), dim3(256), 0, s, comm, psi, pi, out, N);
}

}  // namespace g12
