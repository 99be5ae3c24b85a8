// Device-resident state of one time evolution and the launchers of its kernels.
#pragma once

#include "te_internal.h"

namespace te {

// Scalars of the current Krylov space, written by the Lanczos kernels and consumed by the
// step controller. v_j = s_j u_j; H[j][j] = a_j; H[j+1][j] = -H[j][j+1] = beta_j.
struct KrylovScalars {
    double2 a[kMaxKrylov];
    double beta[kMaxKrylov];
    double s[kMaxKrylov + 1];
    double h;          // last Lanczos norm (core/krylovTimeEvolver.cpp:566-567)
    int broken;        // lucky breakdown in this Krylov step (:551-555)
    int m_hbd;         // Krylov dimension at the breakdown
};

struct LanczosArgs {
    int64_t n;
    const int* row_block;   // stream kernel row blocks (n_row_blocks + 1)
    int n_row_blocks;
    const int* ptr;
    const int* col;
    const double2* val;
    double2* U;        // (m+1) columns of length ld, column-major
    int64_t ld;
    int j;
    int m;
    double2 alpha;     // -i * expFactor (core/krylovTimeEvolver.cpp:31)
    double tolRate;
    KrylovScalars* ks;
    double2* partials;
    unsigned int* ticket;
    const double2* acc_in;   // partial row sums of the preceding column slices (or null)
    VdView vd;               // value-dictionary copy (set by launch_lanczos_apply when used)
    SellView sell;           // its sliced-ELL form (idem)
};

// Persistent control state of timeEvolve (core/krylovTimeEvolver.cpp:261-512), updated by
// the step controller once per Krylov step and read by the host after each step.
struct StepState {
    double t_now, t_sampling, err, t_step;
    double t_now_prev, err_prev;        // values before the current step
    long long index_samples;
    long long sample_base;              // first sample index taken in the current step
    int S;                              // samples taken in the current step
    int n_steps;
    int m;                              // Krylov dimension (shrinks after a breakdown, :358)
    int statusCode;
    int errorCodeFindSubstep;           // persists across steps as in the reference (:298)
    int nbErrRoundOff, nbErrInt;
    int fatal;                          // 1: no small enough step (:203-207); 2: eigen-solver
    int breakdown;                      // lucky breakdown in the current step
    int write_w;                        // the current step updates currentVec (:374-383)
    int pad;
    long long phase_ticks[3];           // controller phases (eigen, step search, samples), wall clock
};

struct StepConsts {
    double t, samplingStep, tolRate, tol, numErrEst;
    long long n_samples;
    int fastIntegration;
    int eig_fast;                       // 1: bisection and twisted factorisations (QL when
                                        //    eigenvalues lie close), 0: QL
    int ts_rows;
    const double2* ts_nodes;            // tanh-sinh (complement, weight), rows concatenated
    const int* ts_off;                  // row offsets, ts_rows + 1 entries
    const double2* gl15;                // Gauss-Legendre (x, w): centre first, then x > 0
    const double2* gl7;
    int gl15_n, gl7_n;
    double2* Y;                         // n_samples x kMaxKrylov sample coefficients
    double2* wcoef;                     // kMaxKrylov coefficients of currentVec
    double* trace;                      // t_step per Krylov step
    long long trace_cap;
};

// Distinct values a diagonal observable may have to be held as one-byte indices.
constexpr int kDiagTab = 256;

struct CombineArgs {
    int64_t n;
    const double2* U;
    int64_t ld;
    int m;
    const double2* wcoef;
    int write_w;                        // U[:,0] = U w   (currentVec, :382)
    double2* U0;                        // == U (column 0), written when write_w
    const double2* Y;                   // sample coefficients, row y_base + s
    long long y_base;
    int S;
    double2* last_out;                  // state of sample y_base + S - 1 when write_last
    int write_last;
    const double* diag;                 // nobs x n, Re of diagonal observables (or null: diag_idx)
    const unsigned char* diag_idx;      // n x idx_stride bytes: row r's indices into diag_tab of
    int idx_stride;                     // the launch's observables (when diag is null)
    const double* diag_tab;             // nobs x kDiagTab distinct values
    int nobs;
    int obs_base;                       // index of the first of these observables in obs_values
    double* obs_values;                 // total_obs x n_samples
    long long n_samples;
    double2* partials;                  // grid x (observables x samples of the launch)
};

// Quadrature tables of the error integrals (tanh-sinh rows, Gauss-Legendre 15 and 7) in device
// memory, computed on the host with libm.
struct DevQuad {
    double2* ts_nodes = nullptr;
    int* ts_off = nullptr;
    double2* gl15 = nullptr;
    double2* gl7 = nullptr;
    int ts_rows = 0, gl15_n = 0, gl7_n = 0;
};
int upload_quadrature(te_ctx* ctx, DevQuad& q);
void free_quadrature(DevQuad& q);
// The StepConsts fields shared by every instance of one parameter set (Y, wcoef and trace are
// left null).
StepConsts step_consts(const te_evolve_params& P, double tolRate, double numErrEst, long long n_samples,
                       const DevQuad& q);
// Final status of a run from its last StepState (core/krylovTimeEvolver.cpp:445-504): the norm
// check of the returned state, then the error counters; logs the reference's warnings.
int final_status(const StepState& h, double numErrEst, double tol, double state_norm);

// ---- adjoint: Krylov coefficients of one backward step ----
constexpr int kGlMax = 48;          // largest Gauss-Legendre order in the node table
constexpr int kAdjMaxNodes = 256;   // quadrature nodes per Krylov step

// What the host reads back after each backward step.
struct AdjStep {
    int n_nodes;         // node states the gradient passes form (n_quad, or fewer: the rank form)
    int n_quad;          // quadrature nodes of the step (order x sub-intervals)
    int m[2];            // Krylov dimension of each basis over the step (after a breakdown: m_hbd)
    int order, n_sub;
    int capped;          // the phase needed more than kAdjMaxNodes nodes (accuracy not assured)
    double phase;        // |omega|_max h: largest phase of the integrand over the step
    double err[2];       // a-posteriori Krylov error integral of each basis over the step
};

struct AdjCoefArgs {
    const KrylovScalars* ks[2];     // basis 0: psi, basis 1: lambda
    int m[2];                       // Krylov dimension
    int m_from_ks[2];               // 1: m_hbd from ks instead when that recurrence broke down
    int from_end[2];                // 0: the basis starts at the step start t_k (coefficients
                                    //    at x = s - t_k); 1: it starts at t_{k+1} (at h - x)
    const double* scale[2];         // norm of the basis' start vector (null: 1)
    double h;                       // step length
    const double2* gl;              // Gauss-Legendre (node, weight) on [-1, 1], order n at gl_off[n]
    const int* gl_off;
    double2* coef[2];               // kAdjMaxNodes x kMaxKrylov coefficients per basis
    double2* coef_end[2];           // kMaxKrylov: the basis vector combination at t_k
    double* w;                      // kAdjMaxNodes quadrature weights (on [t_k, t_{k+1}])
    AdjStep* out;
    int eig_fast;                   // 1: bisection and twisted factorisations (QL when eigenvalues
                                    //    lie close), 0: QL
};
// The step's rule and node coefficients; with rank_tol > 0 the node sum is then replaced by a
// low-rank form of at most as many terms (adjoint_rank_kernel), exact to rank_tol relative.
void launch_adjoint_coef(const AdjCoefArgs& a, double rank_tol, hipStream_t s);

// Per-instance control flags of the batched driver, updated by the batched step controller.
struct BatchCtl {
    int live;       // the instance still has samples to take and no fatal error
    int stepped;    // the controller completed a Krylov step of it in the current round
    int cm;         // Krylov dimension of that step (the columns the combine pass reads)
    int stall;      // consecutive steps with t_step = 0 and no sample
};

void launch_lanczos_apply(LanczosArgs a, const te_matrix* A, hipStream_t s);
void launch_lanczos_project(const LanczosArgs& a, int grid, hipStream_t s);
// Iteration j of two recurrences on the same matrix (a.j == b.j, same n and m; own bases,
// scalars and partials): one pass over the matrix words for both when A is held as
// sliced ELL (grid blocks, a multiple of 8; 0: the single apply's grid), else one apply after
// the other. Results as from two single launches of the same grid.
void launch_lanczos_apply2(LanczosArgs a, LanczosArgs b, const te_matrix* A, int grid, hipStream_t s);
void launch_lanczos_project2(const LanczosArgs& a, const LanczosArgs& b, int grid, hipStream_t s);
void launch_step_control(const StepConsts& c, StepState* st, KrylovScalars* ks, hipStream_t s);
// One block per instance (c, st, ks, bc: n_inst entries each, device memory).
void launch_batch_step_control(const StepConsts* c, StepState* st, KrylovScalars* ks, BatchCtl* bc, int n_inst,
                               long long n_samples, hipStream_t s);
// Picks the sample/observable chunking; S <= max_samples_per_launch(nobs).
int combine_samples_per_launch(int nobs);
int combine_obs_per_launch();
void launch_combine(const CombineArgs& a, int grid, hipStream_t s);
// Samples a pass of the tiled combine (1 to 16 diagonal observables a launch; grid x 128
// partials), or 0 when it does not apply.
int combine_tiled_samples(int nobs);
void launch_combine_tiled(const CombineArgs& a, int grid, hipStream_t s);
void launch_store_obs(const double2* res, int slot, int mode, double* out, hipStream_t s);

}  // namespace te
