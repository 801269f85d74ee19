// Internal declarations shared by the HIP kernels and the host driver.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "../../include/g12kba.h"
#include "sym_units.h"

namespace g12 {

// Scalars of one RK4 stage as the derivative routines see them (built on the
// device from TrajParams + StageShared, or on the host for routine calls).
struct StageParams {
    double c;         // Y = S + c * K_prev   (stage 0: K_prev is zero)
    double wk;        // RK weight 1, 2, 2, 1
    double dt;        // full step, for the final update (A*dt)/6
    int first;        // stage 0: A = w K
    int last;         // stage 3: S += ((A + w K) * dt) / 6
    int corr;         // correlations on (U != 0 and not HF)
    double U;         // effective onsite interaction
    int nV;           // long-range shells
    double V[G12_MAX_SHELLS];
    double wq;        // quench shift on sites < Nq
    int Nq;
    double t1, t2;
    double eps[G12_MAX_NB];
};

// Per-trajectory constants.
struct TrajParams {
    int Ns, Nb, N;
    int corr;          // U != 0 and not HF
    int as_steps;      // length of this trajectory's adiabatic switching
    int Nq;
    int nV;
    double U, t1, t2;
    double qstr;       // quench strength
    double Vprof[G12_MAX_SHELLS];  // V_n = U_eff * Vprof[n-1], or V_n = Vprof[n-1] with v_abs
    int v_abs;         // routine-level calls: V given as such (also with U = 0)
    double eps[G12_MAX_NB];
};

// Per-stage scalars shared by every trajectory of a batch.
struct StageShared {
    double c, wk, dt;
    double cn;         // next stage's input coefficient: Y = S + cn K (stages 0-2)
    int first, last;
    double lam;        // adiabatic switching: U_eff = lam * U (lam = fermi(t)); production: U_eff = U
    double phi;        // quench ramp: shift = phi * qstr on sites < Nq when gate = 1
    int gate;
    int as_phase;      // 1 during adiabatic switching
    int step;          // adiabatic-switching step index (trajectory stops after as_steps)
    int slot;          // trajectory-record slot of this step (production)
    int fold;          // stage 2: the pair-symmetric update folds A into S (k_sym_update)
};

// Where a kernel finds its StageShared: by value, or (graph replay) at
// base[*counter * 4 + stage] in device memory.
struct StageRef {
    StageShared val;
    const StageShared* base;
    const int* counter;
    int stage;
};

// Device buffers of one trajectory.  G2 tensors are stored in the natural
// layout T[p1][p2][q1][q2] (pair index p1*N + p2, one-particle index
// p = site + orb*Ns) on the general path, and in the Hermitian half layout
// (sym_units.h, hblock) on the pair-symmetric path; index 0 is xxxx, 1 is
// xyxy.
struct DevState {
    int Ns, Nb, N;
    int64_t n2, n3, n4;
    double2 *S1, *K1, *A1, *Y1, *h, *I;
    double2* G1t;                      // conj Y1[Y][X] at [X][Y]: the conjugate transpose of the stage G1
    double *Wxy, *Wxx;                 // N x N interaction tables (xy, xx)
    // state, next stage input, RK accumulator, strided-pass result; in
    // mixed-precision runs Y2 and R2 hold complex64 elements
    double2 *S2[2], *Y2[2], *A2[2], *R2[2];
    // mixed-precision runs: the complex64 copy of S (laid out like Y2), the
    // stage-0 input, written with every new state; null in FP64 runs
    double2* Sc[2];
    // pair-symmetric update (k_sym_update): the unit table of this N
    // (su_table, sym_units.h), shared by every trajectory of that N; su_nblk
    // blocks per spin
    const uint64_t* su_unit;
    const uint32_t* su_first;
    int su_nblk;
    double2 *D[2];                     // D[X][Y][Z] = T[X][Z][Y][Z] of the stage input
    double2 *F[2];                     // F[X][Y][Z]
    double2 *Qp[2];                    // Qp[X][Y][Z] = Q[Z][X][Y]
    int* fail;                         // 0 while running, step+1 after the guard trips
    double2* traj;                     // device trajectory ring (traj_slots x N^2) or null
    int traj_slots;
    // collision term J = I + I^dagger of each production step's first stage,
    // ring like traj (correlated trajectories), or null
    double2* jtraj;
    // a given collision term at the step times of the current window
    // ((traj_slots + 1) x N^2) that an uncorrelated (HF) trajectory adds to its
    // G1 equation when jext_on: the learned closure's G1 propagation
    double2* jext;
    int* jext_on;
};

// A set of trajectories advanced together: device arrays of states and
// constants, and (for the tile passes) buckets of trajectories with the same
// padded tile size.
// Table launches: one block per (X, G12_TABLE_GROUP x G12_TABLE_GROUP output
// block) of each correlated trajectory (one block per X at N <= 64).
#define G12_TABLE_GROUP 64
__host__ __device__ inline int table_blocks_per_x(int N) {
    const int nb = (N + G12_TABLE_GROUP - 1) / G12_TABLE_GROUP;
    return nb * nb;
}

struct BatchDev {
    const DevState* ds;      // device array [ntraj]
    const TrajParams* tp;    // device array [ntraj]
    int ntraj;
    int maxN;
    // correlated trajectories and the prefix of their table blocks
    // (N x table_blocks_per_x(N) each), so that a table launch has no idle
    // blocks
    const int* tlist;        // device array [ncorr]
    const int64_t* tprefix;  // device array [ncorr + 1]
    int ncorr;
    int64_t tblocks;
    // pair-symmetric update (k_sym_update): prefix of its blocks over the
    // correlated trajectories (sym_update_blocks each)
    const int64_t* sprefix;  // device array [ncorr + 1]
    int64_t sunits;
    int sym_w8;              // every correlated N is a multiple of 16 (k_sym_update_w8)
};

struct Bucket {
    int NP;                  // padded tile size: N rounded up to a multiple of 16 (wide tiles past 64)
    const int* list;         // device array of trajectory indices
    const int64_t* prefix;   // device array [n + 1]: tiles before each listed trajectory (both spins)
    int n;
    int64_t tiles;
    int waves;               // waves per workgroup, strided pass (autotuned)
    int waves_fused;         // waves per workgroup, fused pass (autotuned)
    int wg_per_cu;           // workgroups per CU cap, strided pass (0: as many as LDS and waves allow)
    int wg_per_cu_fused;     // the same for the fused pass
    int full;                // every listed trajectory has N = NP (no padded tile elements)
    // the pair-symmetric product pass's tile list: tiles (t0, t1), t0 <= t1,
    // in groups of 4 x 4 (kernels.hip ref_of), 2 x 16 hgroups(N) slots per
    // listed trajectory
    const int64_t* hprefix;  // device array [n + 1]
    int64_t htiles;
};

// batched stage kernels (kernels.hip)
void launch_b_stage_setup(const BatchDev& b, const StageRef& sr, int precision, hipStream_t s);
void launch_b_g1_rhs(const BatchDev& b, const StageRef& sr, hipStream_t s);
void launch_b_record(const BatchDev& b, const StageRef& sr, hipStream_t s);
// tile pass of a bucket: strided (fused = 0), fused (1), or with sym = 1 the
// contiguous product pass of the pair-symmetric path
void launch_b_tile(const BatchDev& b, const Bucket& k, const StageRef& sr, int fused, int precision, hipStream_t s,
                   int sym = 0);
// second pass of the pair-symmetric path: C = R + P(R) and the RK update
void launch_b_sym_update(const BatchDev& b, const StageRef& sr, int precision, hipStream_t s);
// flag |= 1 unless the natural-layout tensor T is pair-symmetric (and, with
// herm, Hermitian) within tol
void launch_pair_check(const double2* T, int N, double tol, int herm, int* flag, hipStream_t s);
// largest N of the pair-symmetric path
constexpr int SYM_MAX_N = 64;
// natural <-> Hermitian half layout of one N^4 tensor (out of place; the half
// form's diagonal tiles are made exactly Hermitian from their upper triangles)
void launch_nat_to_half(const double2* nat, double2* half, int N, hipStream_t s);
void launch_half_to_nat(const double2* half, double2* nat, int N, hipStream_t s);
// group pairs (A <= B) of 4 x 4 tiles of the pair-symmetric product pass
__host__ __device__ inline int hgroups(int N) {
    const int ng = (N + 3) / 4;
    return ng * (ng + 1) / 2;
}
// partial traces D of the state and, in mixed-precision runs, its complex64
// copy Sc; half = 1: the state is in the Hermitian half layout
void launch_b_d_from_state(const BatchDev& b, const Bucket& k, int precision, hipStream_t s, int half = 0);
// FP64 tile pass of a bucket with the given products (hmode: 1 +h.X, 2 -X.h),
// sources (0/1) and output mode: the routine-level entries' route through the
// production epilogues (strided pass: R = [O +] sources; fused pass: the RK
// update of its stage)
void launch_b_tile_args(const BatchDev& b, const Bucket& k, const StageRef& r, int fused, int hmode, int sources,
                        int accumulate, hipStream_t s);
void launch_b_advance_counter(int* counter, hipStream_t s);
int device_cu_count();
int default_waves(int NP);
int tile_wave_options(int NP, int* opts);  // compiled wave counts for NP; returns their number

// routine-level helpers (kernels.hip)
void launch_ref_to_nat(const double2* in, double2* out, int Ns, int Nb, hipStream_t s);
void launch_nat_to_ref(const double2* in, double2* out, int Ns, int Nb, hipStream_t s);
// out = [hx (x) 1 + 1 (x) hx] T - T [hy (x) 1 + 1 (x) hy] (natural layout);
// synchronous: 0, G12_E_NOMEM or G12_E_HIP
int launch_comm2(const double2* T, const double2* hx, const double2* hy, double2* out, int N, int precision,
                 hipStream_t s);
// The same enqueued on s without waiting, its launch descriptors written (by a
// kernel) into desc_area, COMM2_DESC_BYTES of device memory that must stay
// untouched until the launches complete; allocates nothing.
constexpr int COMM2_DESC_BYTES = 4096;
int enqueue_comm2(const double2* T, const double2* hx, const double2* hy, double2* out, int N, int precision,
                  void* desc_area, hipStream_t s);
void launch_comm1(const double2* G, const double2* h, double2* out, int N, hipStream_t s);
void launch_pair_diag(const double2* T, double2* out, int N, hipStream_t s, int half = 0);
void launch_g1_step(const double2* comm, const double2* I, double2* K, double2* RK, double w, int N,
                    hipStream_t s);
// out = comm + psi + pi - conj(pi at the particle-swapped transposed element)
// (natural layout): the G2_step sum (G1_G2_t_step.h:55-86) ahead of the fused
// epilogue
void launch_g2_sum(const double2* comm, const double2* psi, const double2* pi, double2* out, int N, hipStream_t s);

}  // namespace g12
