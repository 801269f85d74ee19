// Internal declarations shared by the HIP kernels and the host driver.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/g12kba.h"

namespace g12 {

// Per-stage scalars of one RK4 stage (computed on the host).
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

// Device buffers of one trajectory.  G2 tensors are stored in the natural
// layout T[p1][p2][q1][q2] (pair index p1*N + p2, one-particle index
// p = site + orb*Ns); index 0 is xxxx, 1 is xyxy.
struct DevState {
    int Ns, Nb, N;
    int64_t n2, n3, n4;
    double2 *S1, *K1, *A1, *Y1, *h, *I;
    double *Wxy, *Wxx;                 // N x N interaction tables (xy, xx)
    double2 *S2[2], *K2[2], *A2[2], *R2[2];
    double2 *D[2];                     // D[X][Y][Z] = T[X][Z][Y][Z]
    double2 *F[2], *Ft[2];             // F[X][Y][Z], Ft[Z][Y][X]
    double2 *Q[2], *Qp[2];             // Q[Z][X][Y], Qp[X][Y][Z]
    int* fail;                         // 0 while running, step+1 after the guard trips
    double2* traj;                     // device trajectory ring (traj_slots x N^2)
    int traj_slots;
};

// kernel launchers (kernels.hip)
void launch_stage_g1(const DevState& d, const StageParams& sp, hipStream_t s);
void launch_pass_strided(const DevState& d, const StageParams& sp, int precision, hipStream_t s);
void launch_tables(const DevState& d, const StageParams& sp, hipStream_t s);
void launch_g1_rhs(const DevState& d, const StageParams& sp, hipStream_t s);
void launch_pass_fused(const DevState& d, const StageParams& sp, int precision, hipStream_t s);
void launch_record(const DevState& d, int slot, int64_t step, hipStream_t s);

// routine-level helpers (kernels.hip)
void launch_ref_to_nat(const double2* in, double2* out, int Ns, int Nb, hipStream_t s);
void launch_nat_to_ref(const double2* in, double2* out, int Ns, int Nb, hipStream_t s);
// out = (accumulate ? out : 0) + [hx.X - X.hy] restricted by hmode bits
// (1: left term, 2: right term) over all four pair indices.
void launch_comm2(const double2* T, const double2* hx, const double2* hy, double2* out, double2* tmp,
                  int N, int precision, hipStream_t s);
void launch_hf(double2* h, const double2* G1, int Ns, int Nb, const StageParams& sp, hipStream_t s);
void launch_comm1(const double2* G, const double2* h, double2* out, int N, hipStream_t s);
void launch_d_gather(const double2* T, double2* D, int N, hipStream_t s);
void launch_w_tables(double* Wxy, double* Wxx, int Ns, int Nb, const StageParams& sp, hipStream_t s);
void launch_f_tables(const double2* DP, const double2* DQ, const double* Wxy, const double* Wxx,
                     double2* F, double2* Ft, int N, hipStream_t s);
void launch_q_table(const double2* G1, const double* W, double2* Q, double2* Qp, int N, hipStream_t s);
void launch_collision_from_f(const double2* F, double2* I, int N, hipStream_t s);
void launch_psi(const double2* G1, const double* W, const double2* Q, const double2* Qp, double2* psi,
                int N, hipStream_t s);
void launch_pi(const double2* G1, const double2* F, double2* pi, int N, hipStream_t s);
void launch_g1_step(const double2* comm, const double2* I, double2* K, double2* RK, double w, int N,
                    hipStream_t s);
void launch_g2_step(const double2* comm, const double2* psi, const double2* pi, double2* K, double2* RK,
                    double w, int N, hipStream_t s);

}  // namespace g12
