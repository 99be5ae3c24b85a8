/*
 * timeevolver.h — C-ABI of the MI355X (gfx950) Krylov time-evolution library
 * (libtimeevolver_hip.so). This is the drop-in boundary for TimeEvolver's hot path:
 * the reference's C++ API (core/krylovTimeEvolver.h, core/matrixDataTypes.h,
 * core/krylovObservables.h, helper/Basis.h, helper/hamiltonian.h) is re-implemented by a
 * thin C++ layer on top of these entry points, and any FFI (ctypes, cgo, JNI) can bind
 * them directly. Plain pointers and sizes only; complex numbers are interleaved
 * (re, im) doubles exactly as std::complex<double> lays them out.
 *
 * Error convention: every int-returning function returns TE_OK (0) or a TE_ERR_* code
 * and records a message retrievable with te_last_error(ctx). Where the reference logs
 * FATAL and calls exit(1), the library returns TE_ERR_FATAL instead; the C++ layer turns
 * that back into exit(1). No C++ exception crosses this boundary.
 *
 * Threading: a context owns one HIP stream and is thread-compatible, not thread-safe.
 * All computation runs on the device of the context; host pointers are copied in/out.
 */
#ifndef TIMEEVOLVER_H
#define TIMEEVOLVER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TE_OK 0
#define TE_ERR_ARG 1          /* invalid argument (size, null pointer, dimension mismatch) */
#define TE_ERR_HIP 2          /* a HIP runtime call failed */
#define TE_ERR_ALLOC 3        /* device or pinned-host allocation failed */
#define TE_ERR_FATAL 4        /* the reference would log FATAL and exit(1) */
#define TE_ERR_UNSUPPORTED 5  /* outside what this build supports (e.g. nnz >= 2^31) */
#define TE_ERR_STOP 6         /* a host observable requested a stop (requestStopException) */

typedef struct te_ctx te_ctx;
typedef struct te_matrix te_matrix;
typedef struct te_basis te_basis;
typedef struct te_evolver te_evolver;

/* ---- context ---------------------------------------------------------------------- */
/* No reference counterpart (the reference has no device); owns the HIP stream. */
int te_ctx_create(int device, te_ctx** out);
void te_ctx_destroy(te_ctx* ctx);
const char* te_last_error(const te_ctx* ctx);
int te_ctx_synchronize(te_ctx* ctx);
/* Library version: TIMEEVOLVER_VERSION of core/version.h:7 (201). */
int te_version(void);

/* ---- sparse matrix: replaces TE::smatrix (core/matrixDataTypes.h:59-99) ------------- */
/* smatrix(val, col, row, nbV, nn, mm) + initialize() (core/matrixDataTypes.cpp:108-133,
 * :196-240): zero-based COO with size_t indices (duplicates are kept and summed by the
 * product, as MKL does), converted to device CSR with int32 indices. values: 2*nnz doubles. */
int te_matrix_from_coo(te_ctx* ctx, const double* values, const size_t* col, const size_t* row,
                       size_t nnz, size_t n_rows, size_t n_cols, te_matrix** out);
void te_matrix_destroy(te_matrix* A);
int te_matrix_info(const te_matrix* A, size_t* n_rows, size_t* n_cols, size_t* nnz);
/* No counterpart in the reference: the device storage the evolver's Lanczos products read
 * for A. format: 0 CSR with lanes per row, 1 CSR row blocks, 2 value-dictionary row blocks
 * (32-bit columns and 16-bit indices into n_values distinct off-diagonal values, dense
 * diagonal), 3 the same in a locality order of the rows (lexicographic order of the basis
 * states for operators from te_operator_build, the modes ranked by a Cuthill-McKee numbering of
 * the first such operator's mode couplings), 4 its sliced-ELL form (64-row slices stored
 * slot-major; 4-byte slots of a 16-bit column offset and a 16-bit value index, one base column
 * per slot of a slice; entries at the same column offset and value on every row of a slice as
 * one 8-byte record per slice; the dense diagonal as 8-byte reals when every imaginary part is
 * +0.0, else 16-byte complex values). With format 3, and with 4 in the locality order, te_spmv reads
 * the CSR row blocks in the basis order. stream_bytes: bytes of matrix data one product reads. */
int te_matrix_storage(const te_matrix* A, int* format, size_t* n_values, size_t* stream_bytes);
/* smatrix::norm1 (core/matrixDataTypes.cpp:67-79): maximal absolute column sum. */
int te_matrix_norm1(te_matrix* A, double* out);
/* Copies the device CSR back as COO (rows ascending, columns ascending within a row):
 * the public values/columns/rowIndex arrays of smatrix. */
int te_matrix_to_coo(te_matrix* A, double* values, size_t* col, size_t* row);
/* The device CSR as it is held (the CSR that mkl_sparse_convert_csr builds in the reference,
 * core/matrixDataTypes.cpp:220-221): row_ptr n_rows+1 entries, col and values nnz entries.
 * Any pointer may be NULL to skip that array. */
int te_matrix_to_csr(te_matrix* A, double* values, int64_t* row_ptr, int* col);
/* smatrix::spMV (core/matrixDataTypes.cpp:169-190): out = alpha * A * in, host pointers. */
int te_spmv(te_matrix* A, double alpha_re, double alpha_im, const double* in, double* out);
/* Same on device pointers (n_cols and n_rows complex128 each), asynchronous on the
 * context's stream. */
int te_spmv_device(te_matrix* A, double alpha_re, double alpha_im, const void* in_dev, void* out_dev);
/* No counterpart in the reference: te_spmv for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Same arguments, kernels
 * and results as te_spmv, but it allocates nothing: in and out are staged through a fixed
 * device array of the library (2 x 2,097,152 complex128, 64 MiB, held on every device the
 * library runs on until the process ends). Matrices with more than 2,097,152 rows or columns
 * return TE_ERR_UNSUPPORTED. Every call on a device shares that device's array: calls that
 * overlap on one device (two contexts or threads) race. */
int te_spmv_noalloc(te_matrix* A, double alpha_re, double alpha_im, const double* in, double* out);

/* ---- basis: replaces basis / tensorBasis (helper/Basis.h:277-357) -------------------- */
/* basis(N, K, nbQudits, capacity, quant) (helper/Basis.cpp:37-120), built on the device. */
int te_basis_create(te_ctx* ctx, int N, int K, int nbQudits, int capacity, int quant, te_basis** out);
/* tensorBasis(N1, K1, N2, K2, C2) (helper/Basis.cpp:131-182). */
int te_tensor_basis_create(te_ctx* ctx, int N1, int K1, int N2, int K2, int C2, te_basis** out);
void te_basis_destroy(te_basis* b);
/* numberElements and numberModes (helper/Basis.h:280-282). */
int te_basis_info(const te_basis* b, size_t* n_elements, int* n_modes);
/* basisElements (helper/Basis.h:283): n_elements x n_modes ints, row-major. */
int te_basis_states(const te_basis* b, int* occ);
/* hashTable.find (helper/Basis.h:284): index of each of `count` occupation vectors
 * (count x n_modes ints), or -1 when absent. */
int te_basis_find(const te_basis* b, const int* occ, size_t count, int64_t* index_out);

/* ---- operator builder: replaces Hamiltonian::createMatrix (helper/hamiltonian.cpp:318-401) */
/* One opTerm (helper/hamiltonian.h:18-35): coef * product of operators applied in list
 * order; operations[i] = +1 creation, -1 annihilation on modes[i], capacities[i] = 0 for
 * no maximal occupation. */
typedef struct {
    double coef_re, coef_im;
    int n_ops;
    const int* operations;
    const int* modes;
    const int* capacities;
} te_op_term;
/* createMatrix(terms, basis): bit-identical CSR (rows ascending, columns ascending, entries
 * with |v| < DBL_EPSILON dropped), computed on the device. */
int te_operator_build(te_ctx* ctx, const te_basis* b, const te_op_term* terms, int n_terms, te_matrix** out);

/* ---- time evolution: replaces krylovTimeEvolver (core/krylovTimeEvolver.h:56-133) -------- */
typedef struct {
    double t;              /* total time (:26 t) */
    double sampling_step;  /* :26 samplingStep */
    double exp_factor;     /* :26 expFactor (default 1.0): evolves exp(-i expFactor H t) */
    double tol;            /* :26 tol (default 1e-6) */
    int m;                 /* :26 mm (default 40); the run uses min(m, dimension) (:35), which must
                              be at most 64 (else TE_ERR_UNSUPPORTED) */
    int fast_integration;  /* :26 fastIntegration (default 0) */
} te_evolve_params;

#define TE_OBS_SPARSE 0    /* krylovSpMatrixObservable: Re<psi|O|psi> (core/krylovObservables.cpp:173-194) */
#define TE_OBS_VECTOR 1    /* krylovVectorObservable: |<psi|o>|^2 (core/krylovObservables.cpp:216-236) */
#define TE_OBS_HOST 2      /* any other krylovBasicObservable::expectation, called on the host */

/* Host observable callback: state is n complex128 (host memory, valid during the call).
 * Return 0 and set *value, or nonzero to request a stop (requestStopException,
 * core/krylovObservables.h:16-20). */
typedef int (*te_obs_callback)(void* user, const double* state, size_t n, double* value);

typedef struct {
    int kind;                /* TE_OBS_* */
    te_matrix* matrix;       /* TE_OBS_SPARSE */
    const double* vector;    /* TE_OBS_VECTOR: n complex128, copied at evolver creation */
    te_obs_callback callback;/* TE_OBS_HOST */
    void* user;
} te_observable;

/* krylovReturn (core/krylovTimeEvolver.h:32-51) minus the moved-out objects. */
typedef struct {
    double err;
    double evolved_time;
    size_t num_samples;
    size_t n_steps;
    size_t dim;
    size_t krylov_dim;
    int status_code;
} te_result;

/* Constructor (core/krylovTimeEvolver.cpp:26-84): validates (dimension > 0, |‖psi0‖-1| <= tol,
 * else TE_ERR_FATAL), computes norm1, m = min(m, dim), n_samples = floor(t/step)+1 and uploads
 * psi0. The matrix and observables are borrowed and must outlive the evolver. */
int te_evolver_create(te_ctx* ctx, te_matrix* H, const double* psi0, const te_evolve_params* params,
                      const te_observable* obs, int n_obs, te_evolver** out);
/* Same, with psi0 already in device memory (n complex128). */
int te_evolver_create_device(te_ctx* ctx, te_matrix* H, const void* psi0_dev, const te_evolve_params* params,
                             const te_observable* obs, int n_obs, te_evolver** out);
void te_evolver_destroy(te_evolver* ev);
/* Number of samples per observable (n_samples of :57). */
int te_evolver_num_samples(const te_evolver* ev, size_t* n_samples);
/* timeEvolve() (core/krylovTimeEvolver.cpp:261-512), from psi0 each time it is called.
 * obs_values: n_obs x n_samples doubles (may be NULL); evolved_state: 2*dim doubles, the last
 * sampled state (krylovReturn::evolvedState, :665; may be NULL); result may not be NULL.
 * Returns TE_OK also for soft failures reported through result->status_code. */
int te_evolver_run(te_evolver* ev, double* obs_values, double* evolved_state, te_result* result);
/* krylovBasicObservable::sampleIndex (core/krylovObservables.h:60) of each observable after the
 * last run (counts: n_obs entries): the run's numSamples, plus one for the observables listed
 * before a host observable that requested a stop, which sample() had already evaluated at the
 * stop sample (core/krylovTimeEvolver.cpp:120-126). Samples an observable did not record read 0
 * in te_evolver_run's obs_values. */
int te_evolver_obs_recorded(const te_evolver* ev, size_t* counts);
/* For measurement harnesses (no reference counterpart): bytes per row of diagonal-observable
 * data each combine pass reads. Diagonal observables with at most 256 distinct values (all of
 * them, else none) are held as one-byte indices into tables of those values, 16 bytes per row
 * per 16 observables (TE_DIAG_BYTES=0 in the environment keeps 8-byte diagonals); the values
 * used are the same doubles. */
int te_evolver_diag_bytes(const te_evolver* ev, size_t* bytes_per_row);
/* Device pointer to the last sampled state (valid until the next run/destroy). */
int te_evolver_state_device(te_evolver* ev, const void** state_dev);
/* The Krylov step sizes of the last run (up to cap entries; returns the count in *n). */
int te_evolver_step_trace(te_evolver* ev, double* t_steps, size_t cap, size_t* n);

/* Kernel timing of te_evolver_run, for measurement harnesses (no reference counterpart):
 * HIP events recorded around the first 4096 launches of each kind on the context's stream,
 * for the kinds whose bit is set in kinds (bit k: kind k; 0 turns timing off; events cost a
 * few µs each). kind: 0 Lanczos apply (fused SpMV, the dominant kernel), 1 Lanczos project,
 * 2 step controller, 3 combine (state update, samples, observables); the events of kinds 0, 1
 * and 3 bracket the kernel and the one-block launch that sums its partials; 4..6 the controller's
 * phases (eigen-decomposition, step-size search, sample coefficients) from the device clock,
 * with *launches the number of Krylov steps. */
int te_evolver_set_profiling(te_evolver* ev, int kinds);
int te_evolver_kernel_time(const te_evolver* ev, int kind, double* total_ms, size_t* launches);

/* ---- batched evolution (no reference counterpart: the reference runs one krylovTimeEvolver
 * per instance, core/krylovTimeEvolver.cpp:26-84 and :261-512) ----------------------------- */
typedef struct te_batch te_batch;
/* n_inst independent time evolutions advanced together on the device, every one exactly the
 * plain method of te_evolver_create/te_evolver_run (same Lanczos recurrence, step control,
 * sampling and status rules) with its own Hamiltonian H[i] (square, rows of at most 257
 * entries), initial state psi0[i] (H[i] dimension complex128, host memory, copied) and n_obs
 * observables obs[i * n_obs + o] (diagonal TE_OBS_SPARSE matrices only; n_obs <= 64, read 4 a pass
 * over each instance's Krylov basis). One
 * parameter set for all. Validation per instance as the constructor: TE_ERR_FATAL names the
 * instance whose state is not normalised (:45-48). Matrices are borrowed (must outlive the
 * batch). Instances given the same matrix share its passes: four (or two) of them are served by
 * one pass over its words, with the same result bits (TE_BATCH_SHARED=0: each on its own). */
int te_batch_create(te_ctx* ctx, int n_inst, te_matrix* const* H, const double* const* psi0,
                    const te_evolve_params* params, const te_observable* obs, int n_obs, te_batch** out);
void te_batch_destroy(te_batch* b);
/* Runs every instance from its psi0 to t. obs_values: n_inst x n_obs x n_samples doubles (may be
 * NULL); states: n_inst pointers (the array or any entry may be NULL) receiving each instance's
 * last sampled state; results: n_inst entries. An instance that hits a fatal condition of the
 * reference (no small enough step, :203-207) stops alone with status_code -1 and the call
 * returns TE_ERR_FATAL naming the first such instance; the others' results are complete. */
int te_batch_run(te_batch* b, double* obs_values, double* const* states, te_result* results);
/* Rows of the packed vectors, samples per observable, and rounds of the last run (one round is
 * one Krylov step of every live instance). */
int te_batch_info(const te_batch* b, size_t* packed_rows, size_t* n_samples, long long* rounds);
/* 1 (default): each round is replayed from a captured HIP graph; 0: launched kernel by kernel.
 * The results are identical. A round that cannot be captured or instantiated is launched kernel
 * by kernel (logged as a warning). */
int te_batch_set_graph(te_batch* b, int enable);
/* Launch shape of the batched kernels (no reference counterpart): unit_slices sliced-ELL slices
 * per SpMV unit (a multiple of 4; 0 on the CSR row-block path, whose units are fixed), vec_rows
 * rows per vector unit (projection, combine; a multiple of 512), blocks per CU of the SpMV and of
 * the vector launches. By default the first te_batch_run measures a few candidates (one Lanczos
 * iteration of every instance each: unit_slices in {16, 32, 64, 128}, vec_rows in {1024, 2048,
 * 4096, 8192}, blocks per CU in {4, 8, 16}), keeps the fastest and caches it for the process
 * under the instances' sorted (dimension, nnz) pairs; TE_BATCH_TUNE=0 in the environment keeps
 * the fixed defaults instead. Each round's launches cover the units of the instances still
 * live, cut into smaller pieces as fewer remain (TE_BATCH_PLAN in the environment: units per CU
 * to aim for, default 8, 0 never cuts). Neither the shape nor the cuts change a result bit: every
 * per-instance sum adds partials of 4 slices and of 512 rows in index order. te_batch_set_shape fixes a shape (tuner off); all four
 * arguments 0 turn the tuner back on for the next run. te_batch_shape reports the shape in use,
 * source (0 defaults, 1 measured, 2 the process cache, 3 set by the caller) and the time the
 * measurement took (tune_ms, 0 unless measured); any pointer may be NULL. */
int te_batch_set_shape(te_batch* b, int unit_slices, int vec_rows, int spmv_blocks_per_cu, int vec_blocks_per_cu);
int te_batch_shape(const te_batch* b, int* unit_slices, int* vec_rows, int* spmv_blocks_per_cu,
                   int* vec_blocks_per_cu, int* source, double* tune_ms);
/* Kernel timing of te_batch_run, for measurement harnesses (no reference counterpart): HIP
 * events recorded around the first 8192 launches of each kind on the context's stream, for the
 * kinds whose bit is set in kinds (0 turns timing off). While timing is on, rounds are launched
 * kernel by kernel instead of replayed from the graph (same results). kind: 0 Lanczos apply of
 * every instance with the launch that sums its partials, 1 projection with its sums, 2 step
 * controller (one block per instance), 3 combine (state update, samples, observable sums, live
 * count). total_ms and launches cover the last run; te_batch_kernel_times gives the last run's
 * launches of one kind one by one, in launch order (ms[0 .. min(cap, *launches)); ms may be NULL
 * with cap 0 to ask for the count). Apply and projection launches come m_max to a round. */
int te_batch_set_profiling(te_batch* b, int kinds);
int te_batch_kernel_time(const te_batch* b, int kind, double* total_ms, size_t* launches);
int te_batch_kernel_times(const te_batch* b, int kind, double* ms, size_t cap, size_t* launches);

/* ---- adjoint gradient (no reference counterpart: the reference only evolves forward) ---- */
typedef struct te_adjoint te_adjoint;
typedef struct {
    double forward_ms;       /* device time of the forward evolution */
    double backward_ms;      /* device time of the backward pass */
    int n_steps;             /* Krylov steps of the forward schedule (replayed backward) */
    long long nodes;         /* quadrature nodes over all steps */
    double max_phase;        /* largest integrand phase |omega| h of a step */
    int capped;              /* steps whose rule was capped at the node limit (0 expected) */
    double backward_err;     /* sum of the a-posteriori Krylov error integrals of lambda */
    double psi_err;          /* the same for back-evolved psi (checkpoint interval >= 1) */
    long long node_states;   /* node states formed per basis over all steps: each step's node
                                sum in low-rank form (at most its quadrature node count) */
} te_adjoint_info;
/* Loss L = <psi(t)|O|psi(t)> for psi(t) = exp(-i e H t) psi0 (e = params->exp_factor, evolved by
 * the plain method with params' t, tol, m and fast_integration; the sampling step is ignored)
 * and its gradient with respect to theta_p for H = H_0 + sum_p theta_p G_p:
 *   dL/dtheta_p = 2 e int_0^t Im <lambda(s)|G_p|psi(s)> ds,  lambda(s) = exp(i e H (t-s)) O psi(t).
 * H: the Hamiltonian at the current theta; generators[p] = G_p (n_params <= 64, Hermitian, real
 * diagonal; borrowed); observable O: diagonal. checkpoint_interval K: 0 keeps every forward
 * Krylov basis (no recomputation; memory steps x (m+1) x dim complex128); K >= 1 keeps psi(t_k)
 * every K steps and back-evolves psi with lambda, re-anchored at each checkpoint. */
int te_adjoint_create(te_ctx* ctx, te_matrix* H, int n_params, te_matrix* const* generators, te_matrix* observable,
                      const te_evolve_params* params, int checkpoint_interval, te_adjoint** out);
void te_adjoint_destroy(te_adjoint* a);
/* psi0: dim complex128 (host; |‖psi0‖ - 1| <= tol else TE_ERR_FATAL). loss: L; grad: n_params;
 * result (may be NULL): the forward run's krylovReturn fields; info (may be NULL). */
int te_adjoint_run(te_adjoint* a, const double* psi0, double* loss, double* grad, te_result* result,
                   te_adjoint_info* info);
/* Replaces H (same dimension and storage order: another te_operator_build of the same basis),
 * keeping the generators; for gradient steps or finite differences at new coefficients. */
int te_adjoint_set_hamiltonian(te_adjoint* a, te_matrix* H);
/* psi(t) of the last run (dim complex128, host). */
int te_adjoint_state(te_adjoint* a, double* psiT);

/* Log level filter (krylovLogger::loggingLevel, core/krylovLogger.h:17: DEBUG, INFO, WARNING,
 * ERROR, FATAL = 0..4; default 0 prints everything, as the unfiltered Boost.Log core does).
 * Messages go to stderr. */
void te_set_log_level(int level);

#ifdef __cplusplus
}
#endif
#endif /* TIMEEVOLVER_H */
