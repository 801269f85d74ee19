/*
 * g12kba.h — C ABI of the MI355X-native HF-GKBA G1-G2 (GW) propagator.
 *
 * Drop-in boundary for the reference's RK4 stepper and time-derivative
 * routines (VlcekGroup/G12KBA, headers under G12KBA-GW/src/headers).  Plain pointers and
 * sizes only.  Complex numbers are interleaved (re, im) doubles, exactly the
 * memory of std::complex<double>.  Layouts are the reference's:
 *
 *   N = Ns*Nb, one-particle index X = site + orb*Ns   (HF.h:62-93)
 *   G1[X][Y] at X*N + Y                                 (row-major N x N)
 *   G2[A][B] at A*N*N + B, pair index of p1=(i,a), p2=(j,b):
 *       A = i + j*Ns + (a + b*Nb)*Ns*Ns                 (commutators.h:59-60)
 *
 * Conventions (reference semantics unless noted):
 *   - every routine returns 0 on success or a negative G12_E* code; no
 *     exception crosses the ABI;
 *   - outputs are fully overwritten, except `h` in g12_hf_hamiltonian and
 *     `RK_result` in g12_g1_step / g12_g2_step, which accumulate (+=);
 *   - calls are synchronous at return; a context is used by one host thread;
 *   - one GPU (device 0 unless G12KBA_DEVICE is set).
 */
#ifndef G12KBA_H
#define G12KBA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define G12_OK 0
#define G12_GUARD 1          /* RK4 density guard tripped (reference return value 1) */
#define G12_E_ARG (-1)       /* invalid argument */
#define G12_E_HIP (-2)       /* HIP runtime failure */
#define G12_E_NOMEM (-3)     /* device allocation failed */
#define G12_E_SIZE (-4)      /* size outside what the kernels support */
#define G12_E_IO (-5)        /* file I/O failure */

/* quench protocol */
#define G12_QUENCH_FULL 0
#define G12_QUENCH_PULSE 1

/* long-range interaction law V_n (n = 1..Ns-1) for EHM = true */
#define G12_VLAW_EXP 0       /* V_n = U exp(-decay_rate n)   (RK_solver.h:74, intended law) */
#define G12_VLAW_COULOMB 1   /* V_n = U / (1 + n) */
#define G12_VLAW_EXPLICIT 2  /* V_n = U * V_profile[n-1] */

/* precision of the two-particle contractions ([H2,G2]); G1, I and the RK
 * state are FP64 in every mode */
#define G12_PREC_FP64 0
#define G12_PREC_FP32_3M 1   /* 3M complex FP32 products, FP32 accumulation */
#define G12_PREC_BF16X3 2    /* BF16 hi/lo split (hh + hl + lh), FP32 accumulation */

/* Sizes: N = Ns*Nb <= G12_MAX_N in FP64 (the reference takes any Ns, Nb:
 * read_input.h:85-146, Main.cpp:48-50), N <= G12_MAX_N_MIXED in the
 * mixed-precision modes, whose tile products hold a whole N x N tile in LDS. */
#define G12_MAX_N 128
#define G12_MAX_N_MIXED 64
#define G12_MAX_NB G12_MAX_N
#define G12_MAX_SHELLS 128

/* Run parameters: every global of read_input.h:27-49 plus the choices the
 * reference leaves undefined or hard-codes (its AS dt, tail length, quench
 * gate time, V law, orbital energies, filling). */
typedef struct g12_params {
  int Ns, Nb;
  double t1, t2, t3;          /* t3 is read but unused, as in the reference */
  double U;
  int time_steps;
  double dt;                  /* dt_fixed */
  double quench_strength;
  int Nq;
  int q_type;                 /* G12_QUENCH_* ("none" = full with strength 0, Nq 0) */
  int HF;                     /* 1: Hartree-Fock only (G2 frozen, no collision term) */
  int EHM;                    /* 1: long-range interactions V_n */
  double AS_rate, AS_midpoint;
  double quench_rate, quench_on_midpoint, quench_off_midpoint;
  double decay_rate;          /* 0.7 (read_input.h:29) */
  int V_law;                  /* G12_VLAW_* */
  int nV_profile;             /* entries in V_profile (EXPLICIT law) */
  double V_profile[G12_MAX_SHELLS];
  double eps[G12_MAX_NB];     /* orbital energies, default 0 (reference defect B-4) */
  double AS_dt;               /* 0.1 (AS.h:115) */
  int AS_tail;                /* 20  (AS.h:96) */
  double quench_gate_time;    /* 50  (RK_solver.h:141) */
  int p_number;               /* filled states in G0; <0 means Ns/2 (Main.cpp:52) */
  int precision;              /* G12_PREC_* */
} g12_params;

typedef struct g12_ctx g12_ctx;
typedef struct g12_batch g12_batch;

/* ---- parameters and input file ----------------------------------------- */

/* Compiled defaults of read_input.h:27-49 plus the documented extensions. */
int g12_params_default(g12_params* p);

/* Parse a reference-format `input` file (key line, value on the next line;
 * substring key match, first match wins; a missing key or value gives 0 and a
 * warning on stderr).  Replaces assign_vals, read_input.h:53-146.  Two
 * optional keys extend the format: `long_range_law` (exp | coulomb) and
 * `filling`. */
int g12_parse_input(const char* path, g12_params* p);

/* ---- initial state --------------------------------------------------------
 * G1(0) = i * sum_{k < p_number} phi_k phi_k^dagger over the lowest
 * eigenvectors of the U = 0 Hamiltonian.  Replaces make_G0, init_G.h:57-75
 * (the reference's FP32 LAPACKE_cheev is replaced by an FP64 Jacobi solver;
 * fp32_emulation = 1 rounds h and the eigenvectors through float). */
int g12_make_g0(const g12_params* p, double* G1, int fp32_emulation);

/* Domain-wall initial state (init_G_test, init_G.h:78-86): every orbital of
 * the first int(Ns/2) sites occupied, G1[X][X] = i there and 0 elsewhere. */
int g12_init_g_test(const g12_params* p, double* G1);

/* ---- context, state, drivers --------------------------------------------- */

/* G12_E_SIZE for N = Ns*Nb beyond G12_MAX_N (G12_MAX_N_MIXED in the
 * mixed-precision modes); G12_E_ARG for a
 * non-finite parameter, AS_rate < 0, AS_dt <= 0, a production clock that
 * never reaches time_steps * dt, or a switching ramp that does not reach U
 * exactly (AS.h:96-100) within 10^7 steps. */
int g12_create(const g12_params* p, g12_ctx** out);
int g12_destroy(g12_ctx* ctx);

/* Upload / download (G1, G2xxxx, G2xyxy) in the reference layouts. */
int g12_set_state(g12_ctx* ctx, const double* G1, const double* G2xx, const double* G2xy);
int g12_get_state(g12_ctx* ctx, double* G1, double* G2xx, double* G2xy);

/* Adiabatic switching, replaces Adiabatic_switching (AS.h:63-214).
 * *steps receives the number of RK4 steps taken. */
int g12_adiabatic_switching(g12_ctx* ctx, int* steps);

/* Production RK4, replaces RK4 (RK_solver.h:64-226).  The G1 matrix before
 * every step is written to traj (capacity traj_cap matrices, may be NULL);
 * *steps receives the number of iterations begun (= trajectory length).
 * Returns 0, G12_GUARD when the density guard trips, or a negative code. */
int g12_rk4(g12_ctx* ctx, double* traj, int64_t traj_cap, int64_t* steps);

/* Advance n RK4 steps of the production protocol from the context's current
 * time without recording (benchmarks, checkpoints).  Same guard semantics:
 * when the guard trips, *steps_done counts the iterations begun up to and
 * including the failing one, and the context's time and step count stop at
 * the state it holds (the state before the failing step); later calls return
 * G12_GUARD with nothing done until g12_set_state or g12_load_state. */
int g12_rk4_advance(g12_ctx* ctx, int64_t n, int64_t* steps_done);

/* Advance n production steps from the context's current time, writing the G1
 * matrix before each step to traj (capacity traj_cap matrices, may be NULL).
 * Guard semantics of g12_rk4_advance.  No reference counterpart: the reference's
 * RK4 always starts at t = 0; this resumes a run (checkpoints, extrapolation
 * fallback). */
int g12_rk4_record(g12_ctx* ctx, int64_t n, double* traj, int64_t traj_cap, int64_t* steps_done);

/* Current production time and step count of the context. */
int g12_get_time(g12_ctx* ctx, double* t, int64_t* steps);

/* Sets the production clock (time and step count) from which the next
 * production step runs: the quench protocol of RK_solver.h:141-151 is
 * evaluated at that time.  No reference counterpart (starts a run mid-way,
 * e.g. the G1 propagation of the learned closure at the end of an RK4
 * prefix). */
int g12_set_time(g12_ctx* ctx, double t, int64_t steps);

/* g12_rk4_record that also writes, for every step run, the collision term
 * J = I + I^dagger of the step's first stage (the term G1_step adds to
 * [h, G1], G1_G2_t_step.h:41) to jtraj (capacity traj_cap N x N matrices;
 * NULL: not recorded; G12_E_ARG for an uncorrelated context, which has none).
 * No reference counterpart: training data of the learned closure. */
int g12_rk4_record_collision(g12_ctx* ctx, int64_t n, double* traj, double* jtraj, int64_t traj_cap,
                             int64_t* steps_done);

/* The learned closure's G1 propagation: an uncorrelated context (HF = 1,
 * G2 frozen) adds a given collision term to its G1 equation,
 * K = -i([h, G1] + J(t)), J given at the step times of its next production
 * steps: J[k] at the start of step k (count matrices for count - 1 steps, at
 * most 257).  Stage 0 takes J[k], stages 1 and 2 (J[k] + J[k+1]) / 2, stage 3
 * J[k+1].  The next production call may run at most count - 1 steps
 * (G12_E_ARG beyond) and uses the term up; J = NULL switches it off.
 * G12_E_ARG for a correlated context.  No reference counterpart. */
int g12_set_collision(g12_ctx* ctx, const double* J, int64_t count);

/* Energy of the context's current state, as g12_energy with full U and the
 * quench shift at the context's current time.  Only G1 and the N^2
 * pair-diagonal elements of each G2 tensor leave the device.  No reference
 * counterpart (README.md:61-63 lists energies as future work). */
int g12_state_energy(g12_ctx* ctx, double* E);

/* Per-kernel timing with HIP events on the context's stream.
 * g12_timing_enable(ctx, 1) starts recording; g12_timing_get returns, for the
 * kernel class `which` (0: two-particle commutator pass over strided tiles,
 * 1: fused commutator + right-hand side pass, 2: all other kernels), the number
 * of launches and their summed duration in milliseconds. */
int g12_timing_enable(g12_ctx* ctx, int on);
int g12_timing_get(g12_ctx* ctx, int which, int64_t* launches, double* total_ms);

/* Execution options of a context: use_graphs = 1 replays each RK4 step as one
 * captured hipGraph (default; timing mode always launches directly),
 * autotune = 1 times the compiled wave counts of the tile passes on the
 * context's own tiles before its first run (default).  A negative value leaves
 * an option unchanged.  No reference counterpart (launch configuration). */
int g12_configure(g12_ctx* ctx, int use_graphs, int autotune);

/* *on = 1 when the context runs the pair-symmetric path: its correlated state
 * is symmetric under the exchange of both pairs and Hermitian as an N^2 x N^2
 * matrix, with G1 anti-Hermitian (every state evolved from G2 = 0 is; checked
 * to 1e-12 of the largest element whenever a state is set or loaded), and
 * N <= 64.  The G2 right-hand side is then the contiguous-tile half of the
 * commutator and sources plus its pair exchange, and its tiles (t1, t0) are
 * the negated conjugate transposes of the tiles (t0, t1): one tile-product
 * pass over the tiles t0 <= t1 instead of two passes over all tiles.  The path
 * keeps both symmetries exactly.  Otherwise (or with G12KBA_PAIR_SYMMETRY=0 in
 * the environment at creation) the general two-pass path runs.  No reference
 * counterpart: the reference evaluates all four commutator terms
 * (commutators.h:53-99). */
int g12_pair_symmetric(g12_ctx* ctx, int* on);

/* Waits for all work on the library's GPU (device-wide synchronisation). */
int g12_synchronize(void);

/* Device footprint of a context in bytes. */
int g12_memory_bytes(g12_ctx* ctx, int64_t* bytes);

/* ---- runs of many trajectories -------------------------------------------
 * A batch advances ntraj independent trajectories together: every RK4 stage is
 * one set of launches over all of them (the tile passes are persistent
 * launches over the tiles of every trajectory with the same padded size).
 * Each trajectory has its own parameter block (sizes, hoppings, U, long-range law,
 * quench strength and extent, HF flag); the time protocol (dt, time_steps,
 * AS_*, quench_rate/midpoints/q_type/gate) and the precision must agree, else
 * G12_E_ARG.  The reference runs one trajectory per process (submit.sh sweeps
 * U by launching processes); these calls replace that loop.
 *
 * g12_batch_adiabatic_switching: steps[i] = switching length of trajectory i
 * (it depends on U through the exact U_AS == U test of AS.h:96-100).
 * g12_batch_rk4: traj[i] (or NULL) receives trajectory i's G1 before every
 * step (capacity traj_cap matrices); steps[i] = its length.  A trajectory
 * whose density guard trips stops there while the others go on; the call
 * returns G12_GUARD if any tripped.  g12_batch_guard copies the guard flags
 * (0, or failing step + 1).  g12_batch_buckets reports, per padded tile size,
 * the number of trajectories and the wave counts in use; it returns the number
 * of buckets. */
int g12_batch_create(const g12_params* p, int ntraj, g12_batch** out);
int g12_batch_destroy(g12_batch* b);
int g12_batch_size(g12_batch* b, int* ntraj);
int g12_batch_memory_bytes(g12_batch* b, int64_t* bytes);
int g12_batch_set_state(g12_batch* b, int i, const double* G1, const double* G2xx, const double* G2xy);
int g12_batch_get_state(g12_batch* b, int i, double* G1, double* G2xx, double* G2xy);
int g12_batch_adiabatic_switching(g12_batch* b, int* steps);
int g12_batch_rk4(g12_batch* b, double* const* traj, int64_t traj_cap, int64_t* steps);
int g12_batch_rk4_advance(g12_batch* b, int64_t n);
/* n more production steps from the batch's current time; traj[i] (or NULL)
 * receives trajectory i's G1 before each of them (capacity traj_cap). */
int g12_batch_rk4_record(g12_batch* b, int64_t n, double* const* traj, int64_t traj_cap);
/* g12_batch_rk4_record that also writes each correlated trajectory's
 * collision terms (as g12_rk4_record_collision) to jtraj[i] (or NULL). */
int g12_batch_rk4_record_collision(g12_batch* b, int64_t n, double* const* traj, double* const* jtraj,
                                   int64_t traj_cap);
int g12_batch_guard(g12_batch* b, int* fail);
int g12_batch_timing_enable(g12_batch* b, int on);
int g12_batch_timing_get(g12_batch* b, int which, int64_t* launches, double* total_ms);
int g12_batch_configure(g12_batch* b, int use_graphs, int autotune);
int g12_batch_tune(g12_batch* b);
int g12_batch_pair_symmetric(g12_batch* b, int* on);
int g12_batch_buckets(g12_batch* b, int cap, int* NP, int* count, int* waves_strided, int* waves_fused);
/* g12_state_energy of trajectory i. */
int g12_batch_state_energy(g12_batch* b, int i, double* E);

/* ---- learned G1 extrapolator ------------------------------------------------
 * No reference counterpart (the reference's README names extrapolation of the
 * dynamics as a direction; nothing in its source implements one).  A temporal
 * MLP applied to every upper-triangle element of G1: from `hist` steps of the
 * element and the quench drive it predicts the next `horizon` steps.  Four
 * layers (tanh, `hidden` wide) plus a linear skip path; BF16 MFMA products
 * with FP32 accumulation, FP32 master weights, Adam.  Features and the exact
 * network are described in g12kba_amd/csrc/extrap.hip.
 *
 * Training data: `n_series` complex64 series of length T (interleaved re/im),
 * the drive coefficient of each series (quench strength x (m_p - m_q)) and the
 * shared drive profile (T values); minibatches are drawn on the device.
 * g12_xp_loss_grad / g12_xp_forward / g12_xp_features expose one pass on given
 * data (gradient and feature checks).  g12_xp_rollout predicts nblocks blocks
 * from hist matrices, feeding predictions back, and reports per block the
 * largest trace drift (against tr_ref, the conserved trace, or the last history
 * matrix's when NULL) and anti-hermiticity error of the network's output;
 * project = 1 then restores Re G_ii = 0 and the conserved Im tr G in each
 * block before feeding it back. */
typedef struct g12_xp g12_xp;
typedef struct g12_xp_config {
  int hist;        /* history steps (64) */
  int horizon;     /* predicted steps per block (16) */
  int hidden;      /* hidden width (320) */
  double dt;       /* time step of the trajectories (0.01) */
  uint64_t seed;   /* weight initialisation and minibatch draws */
  double noise;    /* training only: standard deviation of the Gaussian noise added to the
                      normalised history features of each minibatch (0: none), which makes
                      long autoregressive rollouts less sensitive to their own errors */
} g12_xp_config;

int g12_xp_default_config(g12_xp_config* c);
int g12_xp_create(const g12_xp_config* c, g12_xp** out);
int g12_xp_destroy(g12_xp* x);
int g12_xp_param_count(g12_xp* x, int64_t* n);
int g12_xp_get_params(g12_xp* x, float* params);
int g12_xp_set_params(g12_xp* x, const float* params);
int g12_xp_save(g12_xp* x, const char* path);
int g12_xp_load(g12_xp* x, const char* path);
int g12_xp_set_data(g12_xp* x, const float* series, const float* coef, const float* drive, int64_t n_series, int T);
int g12_xp_train(g12_xp* x, int steps, int batch, double lr, float* loss);
int g12_xp_loss_grad(g12_xp* x, const float* X, const float* T, int M, double* loss, float* grad);  /* loss: r^2 up to |r| = 1, 2|r| - 1 beyond, averaged */
int g12_xp_forward(g12_xp* x, const float* X, int M, float* Y);
int g12_xp_features(g12_xp* x, const double* hist, int N, int Ns, int Nq, const double* w, float* X, double* ref,
                    double* scale);
int g12_xp_rollout(g12_xp* x, const double* hist, int N, int Ns, int Nq, const double* w, const double* tr_ref,
                   int project, int nblocks, double* out, double* checks);
/* The same for a Hermitian series (the collision term J = I + I^dagger of
 * the learned closure): the lower triangle mirrors the upper as J^dagger = J;
 * checks report the largest |tr J - tr_ref| and hermiticity error
 * |J - J^dagger|; project restores Im J_ii = 0 and Re tr J = Re tr_ref. */
int g12_xp_rollout_hermitian(g12_xp* x, const double* hist, int N, int Ns, int Nq, const double* w,
                             const double* tr_ref, int project, int nblocks, double* out, double* checks);

/* ---- routine-level entry points ------------------------------------------
 * One entry per reference routine, same argument roles and order, host
 * buffers in the reference layouts.  The work runs on the GPU. */

/* h += h_HF[G1x] (h_HFx_quench, HF.h:57-104).  G1y is ignored as in the
 * reference; V has nV entries; eps has Nb entries (NULL = zeros). */
int g12_hf_hamiltonian(double* h, const double* G1x, const double* G1y, int Ns, int Nb, double U,
                       double t1, double t2, double w, const double* eps, const double* V, int nV,
                       int quench_extent);

/* comm = h G - G h (hG_commutator, commutators.h:28-51); G, h are N x N. */
int g12_h1_commutator(const double* G, const double* h, double* comm, int N);

/* comm = [h1x (x) 1 + 1 (x) h1x] G2 - G2 [h1y (x) 1 + 1 (x) h1y]
 * (h2xyG2xyxy_commutator, commutators.h:53-99). */
int g12_h2_commutator(const double* G2, const double* h1x, const double* h1y, double* comm, int Ns,
                      int Nb, int precision);

/* No reference counterpart: a variant of g12_h2_commutator for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  G2, h1x, h1y and comm are DEVICE pointers
 * (reference layouts, as g12_h2_commutator); `stream` is a hipStream_t (NULL:
 * the null stream).  The work is enqueued on the stream and the call returns
 * without waiting; it calls no hipMalloc/hipFree and no host copy, so it can
 * be captured in a graph.  Its scratch (launch descriptors and the
 * natural-layout copies of G2 and of the result) is one static device array of
 * G12_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, which limits
 * N = Ns*Nb to G12_NOALLOC_MAX_N (G12_E_SIZE beyond).  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (on different
 * streams, or with the scratch of an earlier call still in flight elsewhere)
 * race.  The call clears the calling thread's HIP last-error state before its
 * launches and reports only their errors. */
#define G12_NOALLOC_MAX_N 24
#define G12_NOALLOC_SCRATCH_BYTES (4096 + 2LL * 24 * 24 * 24 * 24 * 16)
int g12_h2_commutator_noalloc(const double* G2, const double* h1x, const double* h1y, double* comm, int Ns,
                              int Nb, int precision, void* stream);

/* Collision integral (collision_int, collision_integral.h:27-69); note the
 * reference's argument order (xyxy, xxxx). */
int g12_collision(double* Ix, const double* G2xyxy, const double* G2xxxx, int Ns, int Nb, double U,
                  const double* V, int nV);

/* GW two-particle sources (Psi_xxxx psi.h:157-277, Psi_xyxy psi.h:26-154). */
int g12_psi_xxxx(double* psi, const double* G1, int Ns, int Nb, double U, const double* V, int nV);
int g12_psi_xyxy(double* psi, const double* G1, int Ns, int Nb, double U, const double* V, int nV);

/* GW polarisation term (Pi_xyxy, pi.h:25-95); called with the G2 pair in
 * either order exactly as the reference (RK_solver.h:163-164). */
int g12_pi(double* pi, const double* G1y, const double* G2xyxy, const double* G2xxxx, double U,
           int Ns, int Nb, const double* V, int nV);

/* Stage derivative and RK accumulation (G1_step / G2_step,
 * G1_G2_t_step.h:24-47 / 55-86). */
int g12_g1_step(const double* comm, const double* I, double* K, double* RK_result, double w, int Ns,
                int Nb);
int g12_g2_step(const double* comm, const double* psi, const double* pi, double* K,
                double* RK_result, double w, int Ns, int Nb);

/* ---- host helpers -------------------------------------------------------- */

/* Logistic ramp (fermi, fermi_func.h:21-24). */
double g12_fermi(double t, double t0, double tau);

/* Dipole trajectory (dipole, dipole.h:25-33); traj holds T N x N matrices. */
int g12_dipole(double* p, const double* traj, int64_t T, int N);

/* Energy of a state in the reference layouts, summed over both spins:
 * E[0] one-body (hopping, eps, quench shift w on sites < Nq), E[1] Hartree,
 * E[2] exchange, E[3] correlation (from the pair-diagonal elements of G2;
 * 0 when G2xx or G2xy is NULL).  lam scales U and every V_n (1 after
 * adiabatic switching).  The correlation weight is the one the equations of
 * motion conserve: for Nb = 1 and a time-independent h0 the total is constant
 * along a run.  No reference counterpart (README.md:61-63 lists energies as
 * future work); host code, no GPU needed. */
int g12_energy(const g12_params* p, double lam, double w, const double* G1, const double* G2xx, const double* G2xy,
               double* E);

/* Byte-compatible text outputs (write_to_file.h:26-69): <dir>/real_G1.txt,
 * <dir>/imag_G1.txt (one row per matrix element, one column per step, a space
 * after every value) and <dir>/dipole.txt.  `progress` (percent) is written to
 * <dir>/progress.txt when >= 0. */
int g12_write_outputs(const char* dir, const double* traj, int64_t T, int N, const double* dipole,
                      double progress);

/* Binary full-precision state dump / load: (t, steps, G1, G2xx, G2xy) in the
 * reference layouts. */
int g12_save_state(g12_ctx* ctx, const char* path);
int g12_load_state(g12_ctx* ctx, const char* path);
/* The same for trajectory i of a batch.  The trajectories of a batch share
 * one clock: loading sets the batch's time and step count to the file's.  No
 * reference counterpart (the reference runs one trajectory per process). */
int g12_batch_save_state(g12_batch* b, int i, const char* path);
int g12_batch_load_state(g12_batch* b, int i, const char* path);

/* Library version string. */
const char* g12_version(void);

#ifdef __cplusplus
}
#endif

#endif /* G12KBA_H */
