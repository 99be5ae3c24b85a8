// C-ABI implementation: contexts, device CSR matrices and the timeEvolve orchestration.
//
// The host side keeps the reference's control structure (core/krylovTimeEvolver.cpp:261-512)
// but every vector operation and all Krylov-space arithmetic run on the device: per Krylov
// step the host enqueues m Lanczos apply/project pairs, the step controller, then reads one
// small StepState record to learn how many samples the step produced, and enqueues the
// combine pass that builds the new state, the sampled states and the observables.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "te_internal.h"
#include "te_krylov.h"

namespace te {

int g_log_level = 0;

void log_msg(int level, const char* fmt, ...) {
    if (level < g_log_level) return;
    static const char* names[] = {"debug", "info", "warning", "error", "fatal"};
    fprintf(stderr, "[timeevolver %s] ", names[level < 0 ? 0 : (level > 4 ? 4 : level)]);
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fprintf(stderr, "\n");
}

int fail(te_ctx* ctx, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

int stream_grid(const te_ctx* ctx, int64_t work_items) {
    static const int per_cu = [] {
        const char* v = std::getenv("TE_STREAM_BLOCKS_PER_CU");      // launch-shape sweeps
        return v ? std::max(1, std::atoi(v)) : 4;
    }();
    const int cap = std::max(8, ctx->num_cus * per_cu);
    int64_t g = (work_items + kBlock - 1) / kBlock;
    g = std::max<int64_t>(8, std::min<int64_t>(g, cap));
    return (int)((g + 7) / 8 * 8);
}

int project_grid(const te_ctx* ctx, int64_t rows) {
    // Two blocks per CU sweeping one front of 1024-row chunks: on the L = N = 14 ring the
    // projection ran 177 us at 2 against 202 us at 1 and 208 us at 4 blocks per CU.
    static const int per_cu = [] {
        const char* v = std::getenv("TE_PROJECT_BLOCKS_PER_CU");    // launch-shape sweeps
        return v ? std::max(1, std::atoi(v)) : 2;
    }();
    const int cap = std::max(8, ctx->num_cus * per_cu);
    int64_t g = (rows + kBlock - 1) / kBlock;
    g = std::max<int64_t>(8, std::min<int64_t>(g, cap));
    return (int)((g + 7) / 8 * 8);
}

int combine_grid(const te_ctx* ctx, int64_t rows) {
    // Two blocks per CU on one front of kBlock-row tiles: on the L = N = 14 ring (40 basis
    // columns, 14 observables) the combine ran 37.5-38.3 ms a run at 2 against 40.1 at 4 and
    // 41-42 at 1 block per CU.
    static const int per_cu = [] {
        const char* v = std::getenv("TE_COMBINE_BLOCKS_PER_CU");    // launch-shape sweeps
        return v ? std::max(1, std::atoi(v)) : 2;
    }();
    const int cap = std::max(8, ctx->num_cus * per_cu);
    int64_t g = (rows + kBlock - 1) / kBlock;
    g = std::max<int64_t>(8, std::min<int64_t>(g, cap));
    return (int)((g + 7) / 8 * 8);
}

SpmvShape choose_shape(const te_ctx* ctx, int64_t n_rows, int64_t nnz) {
    SpmvShape s;
    const double mean = n_rows > 0 ? (double)nnz / (double)n_rows : 1.0;
    int lpr = 1;
    while (lpr < 32 && lpr < mean) lpr <<= 1;
    s.lanes_per_row = lpr;
    s.grid = stream_grid(ctx, n_rows * lpr);
    return s;
}

// ---- quadrature tables (computed on the host with libm, uploaded once per evolver) ----
struct QuadTables {
    std::vector<double2> ts_nodes;
    std::vector<int> ts_off;
    std::vector<double2> gl15, gl7;
};

// Tanh-sinh nodes t = 0, 1, 2, ... (row 0, h = 1) and odd multiples of 2^-k (row k), each as
// (1 - tanh(pi/2 sinh t), (pi/2) cosh t / cosh^2(pi/2 sinh t)), kept while the complement is
// at least 4 * DBL_MIN (Takahasi & Mori's double-exponential rule, Boost's default 15 rows).
static void build_tanh_sinh(QuadTables& q, int rows) {
    const double half_pi = 1.5707963267948966;
    const double min_c = 4.0 * DBL_MIN;
    auto node = [&](double t) {
        const double u = half_pi * std::sinh(t);
        const double ch = std::cosh(u);
        return make_double2(1.0 / (std::exp(u) * ch), half_pi * std::cosh(t) / (ch * ch));
    };
    q.ts_off.assign(1, 0);
    for (int t = 0;; t++) {
        double2 nd = node((double)t);
        if (!(nd.x >= min_c)) break;
        q.ts_nodes.push_back(nd);
    }
    q.ts_off.push_back((int)q.ts_nodes.size());
    for (int k = 1; k < rows; k++) {
        const double h = std::ldexp(1.0, -k);
        for (long j = 0;; j++) {
            double2 nd = node((2 * j + 1) * h);
            if (!(nd.x >= min_c)) break;
            q.ts_nodes.push_back(nd);
        }
        q.ts_off.push_back((int)q.ts_nodes.size());
    }
}

// Gauss-Legendre nodes of order n by Newton's method on P_n; returns the centre (odd n)
// followed by the positive nodes in increasing order, as (x, w).
static std::vector<double2> gauss_legendre(int n) {
    std::vector<double2> pos;
    for (int i = 1; i <= n; i++) {
        double z = std::cos(M_PI * (i - 0.25) / (n + 0.5));
        double p0, p1, dp = 0.0;
        for (int it = 0; it < 100; it++) {
            p0 = 1.0;
            p1 = z;
            for (int k = 2; k <= n; k++) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-17) break;
        }
        p0 = 1.0;
        p1 = z;
        for (int k = 2; k <= n; k++) {
            const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        if (z >= -1e-300) pos.push_back(make_double2(std::fabs(z) < 1e-15 ? 0.0 : z, 2.0 / ((1.0 - z * z) * dp * dp)));
    }
    std::sort(pos.begin(), pos.end(), [](const double2& a, const double2& b) { return a.x < b.x; });
    return pos;
}

int upload_quadrature(te_ctx* ctx, DevQuad& d) {
    QuadTables q;
    build_tanh_sinh(q, 15);
    q.gl15 = gauss_legendre(15);
    q.gl7 = gauss_legendre(7);
    d.ts_rows = (int)q.ts_off.size() - 1;
    d.gl15_n = (int)q.gl15.size();
    d.gl7_n = (int)q.gl7.size();
    TE_HIP(ctx, hipMalloc(&d.ts_nodes, sizeof(double2) * q.ts_nodes.size()));
    TE_HIP(ctx, hipMalloc(&d.ts_off, sizeof(int) * q.ts_off.size()));
    TE_HIP(ctx, hipMalloc(&d.gl15, sizeof(double2) * q.gl15.size()));
    TE_HIP(ctx, hipMalloc(&d.gl7, sizeof(double2) * q.gl7.size()));
    TE_HIP(ctx, hipMemcpy(d.ts_nodes, q.ts_nodes.data(), sizeof(double2) * q.ts_nodes.size(), hipMemcpyHostToDevice));
    TE_HIP(ctx, hipMemcpy(d.ts_off, q.ts_off.data(), sizeof(int) * q.ts_off.size(), hipMemcpyHostToDevice));
    TE_HIP(ctx, hipMemcpy(d.gl15, q.gl15.data(), sizeof(double2) * q.gl15.size(), hipMemcpyHostToDevice));
    TE_HIP(ctx, hipMemcpy(d.gl7, q.gl7.data(), sizeof(double2) * q.gl7.size(), hipMemcpyHostToDevice));
    return TE_OK;
}

void free_quadrature(DevQuad& d) {
    if (d.ts_nodes) hipFree(d.ts_nodes);
    if (d.ts_off) hipFree(d.ts_off);
    if (d.gl15) hipFree(d.gl15);
    if (d.gl7) hipFree(d.gl7);
    d = DevQuad{};
}

StepConsts step_consts(const te_evolve_params& P, double tolRate, double numErrEst, long long n_samples,
                       const DevQuad& q) {
    StepConsts C{};
    C.t = P.t;
    C.samplingStep = P.sampling_step;
    C.tolRate = tolRate;
    C.tol = P.tol;
    C.numErrEst = numErrEst;
    C.n_samples = n_samples;
    C.fastIntegration = P.fast_integration;
    // eigen-decompositions by bisection and twisted factorisations, QL when eigenvalues lie
    // close (TE_EIG=ql: QL always; read when an evolver or batch is created)
    const char* eig = std::getenv("TE_EIG");
    C.eig_fast = (eig && std::strcmp(eig, "ql") == 0) ? 0 : 1;
    C.ts_rows = q.ts_rows;
    C.ts_nodes = q.ts_nodes;
    C.ts_off = q.ts_off;
    C.gl15 = q.gl15;
    C.gl7 = q.gl7;
    C.gl15_n = q.gl15_n;
    C.gl7_n = q.gl7_n;
    return C;
}

int final_status(const StepState& h, double numErrEst, double tol, double state_norm) {
    int statusCode = h.statusCode;
    int nbErrors = 0;
    if (std::fabs(state_norm - 1.0) > tol) {                                            // :446-452
        log_msg(2, "CRITICAL WARNING: Norm of state vector is not inside specified tolerance. THE DESIRED ERROR "
                   "BOUND WILL LIKELY BE VIOLATED.");
        nbErrors++;
        statusCode = 30;
    }
    const double total = numErrEst * (double)h.n_steps;                                // :459-504
    if (h.nbErrRoundOff != 0) {
        log_msg(2, "CRITICAL WARNING: The computed error bound was smaller than the estimate of the numerical "
                   "error in %d Krylov spaces. THE DESIRED ERROR BOUND WILL LIKELY BE VIOLATED.",
                h.nbErrRoundOff);
        nbErrors++;
        statusCode = 10;
    }
    if (h.nbErrInt != 0) {
        log_msg(2, "CRITICAL WARNING: The error of numerical integration did not meet the required accuarcy in "
                   "%d Krylov spaces. THE DESIRED ERROR BOUND MAY BE VIOLATED.",
                h.nbErrInt);
        nbErrors++;
        statusCode = 20;
    }
    if (total > h.err && h.nbErrRoundOff == 0) {
        if (total > tol) {
            statusCode = 11;
            nbErrors++;
        } else if (statusCode != 1) {
            statusCode = 2;
        }
    }
    if (nbErrors > 1) statusCode = 100;
    return statusCode;
}

}  // namespace te

using namespace te;

// =========================================================================================
// context
// =========================================================================================
extern "C" int te_version(void) { return 201; }

extern "C" void te_set_log_level(int level) { te::g_log_level = level; }

namespace te {
// Staging of te_spmv_noalloc, which calls no allocator: x in the first kNoAllocRows entries, y in
// the second (2 x 32 MiB). Every device the library runs on holds its own copy until the process
// ends; calls on one device share it (and race when they overlap). Host code reaches it only
// through the per-device address that noalloc_setup reads with hipGetSymbolAddress: the array's
// name in host code is a stand-in, not a device address.
constexpr int64_t kNoAllocRows = int64_t(1) << 21;
static __device__ __align__(256) double2 scratch[2 * kNoAllocRows];
static std::mutex g_scratch_mu;
static std::vector<double2*> g_scratch_dev;     // per device: scratch's address there (null: not set up)

// Once per device, with that device current (te_ctx_create): records scratch's address.
static int noalloc_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_dev.size() <= device) g_scratch_dev.resize(device + 1, nullptr);
    if (g_scratch_dev[device]) return TE_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return TE_ERR_HIP;
    }
    g_scratch_dev[device] = static_cast<double2*>(p);
    return TE_OK;
}

static double2* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device < (int)g_scratch_dev.size() ? g_scratch_dev[device] : nullptr;
}
}  // namespace te

extern "C" int te_ctx_create(int device, te_ctx** out) {
    if (!out) return TE_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return TE_ERR_HIP;
    if (device < 0 || device >= n) return TE_ERR_ARG;
    te_ctx* c = new te_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return TE_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->num_cus = prop.multiProcessorCount;
        c->max_lds = (int)prop.sharedMemPerBlock;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return TE_ERR_HIP;
    }
    // te_spmv_noalloc's staging on this device; without it that entry point reports an error
    if (te::noalloc_setup(device) != TE_OK)
        te::log_msg(2, "device %d: the staging array of te_spmv_noalloc has no device address", device);
    *out = c;
    return TE_OK;
}

extern "C" void te_ctx_destroy(te_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char* te_last_error(const te_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int te_ctx_synchronize(te_ctx* ctx) {
    if (!ctx) return TE_ERR_ARG;
    hipSetDevice(ctx->device);
    TE_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TE_OK;
}

// =========================================================================================
// matrices
// =========================================================================================
namespace te {

// Uploads host CSR arrays (int32) and derives the diagonal form when every entry is diagonal.
int matrix_upload(te_ctx* ctx, te_matrix* A, const std::vector<int>& ptr, const std::vector<int>& col,
                  const std::vector<double2>& val) {
    TE_HIP(ctx, hipMalloc(&A->row_ptr, sizeof(int) * ptr.size()));
    TE_HIP(ctx, hipMalloc(&A->col, sizeof(int) * std::max<size_t>(col.size(), 1)));
    TE_HIP(ctx, hipMalloc(&A->val, sizeof(double2) * std::max<size_t>(val.size(), 1)));
    TE_HIP(ctx, hipMemcpy(A->row_ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice));
    if (!col.empty()) {
        TE_HIP(ctx, hipMemcpy(A->col, col.data(), sizeof(int) * col.size(), hipMemcpyHostToDevice));
        TE_HIP(ctx, hipMemcpy(A->val, val.data(), sizeof(double2) * val.size(), hipMemcpyHostToDevice));
    }
    bool diag = A->n_rows == A->n_cols;
    for (int64_t r = 0; diag && r < A->n_rows; r++)
        for (int k = ptr[r]; k < ptr[r + 1]; k++)
            if (col[k] != r) {
                diag = false;
                break;
            }
    A->diagonal = diag;
    if (diag) {
        std::vector<double> d(A->n_rows, 0.0);
        for (int64_t r = 0; r < A->n_rows; r++)
            for (int k = ptr[r]; k < ptr[r + 1]; k++) d[r] += val[k].x;
        TE_HIP(ctx, hipMalloc(&A->diag_re, sizeof(double) * std::max<int64_t>(A->n_rows, 1)));
        TE_HIP(ctx, hipMemcpy(A->diag_re, d.data(), sizeof(double) * A->n_rows, hipMemcpyHostToDevice));
    }
    return prepare_spmv(ctx, A);
}

}  // namespace te

extern "C" int te_matrix_from_coo(te_ctx* ctx, const double* values, const size_t* col, const size_t* row,
                                  size_t nnz, size_t n_rows, size_t n_cols, te_matrix** out) {
    if (!ctx || !out) return TE_ERR_ARG;
    *out = nullptr;
    if (n_rows == 0 || n_cols == 0) return fail(ctx, TE_ERR_FATAL, "Empty matrices are not supported.");
    if (nnz > 0 && (!values || !col || !row)) return fail(ctx, TE_ERR_ARG, "null COO array");
    if (nnz >= (size_t)INT32_MAX || n_rows >= (size_t)INT32_MAX || n_cols >= (size_t)INT32_MAX)
        return fail(ctx, TE_ERR_UNSUPPORTED, "matrix exceeds int32 CSR indexing");
    hipSetDevice(ctx->device);
    for (size_t i = 0; i < nnz; i++)
        if (row[i] >= n_rows || col[i] >= n_cols) return fail(ctx, TE_ERR_ARG, "COO index out of range at %zu", i);
    // norm1 in COO order (core/matrixDataTypes.cpp:67-79)
    std::vector<double> colVal(n_cols, 0.0);
    for (size_t i = 0; i < nnz; i++) colVal[col[i]] += std::hypot(values[2 * i], values[2 * i + 1]);
    // CSR ordered by (row, col); stable for duplicates
    std::vector<size_t> order(nnz);
    std::iota(order.begin(), order.end(), (size_t)0);
    bool sorted = true;
    for (size_t i = 1; i < nnz && sorted; i++)
        if (row[i] < row[i - 1] || (row[i] == row[i - 1] && col[i] < col[i - 1])) sorted = false;
    if (!sorted)
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
            return row[a] != row[b] ? row[a] < row[b] : col[a] < col[b];
        });
    std::vector<int> ptr(n_rows + 1, 0), c(nnz);
    std::vector<double2> v(nnz);
    for (size_t i = 0; i < nnz; i++) {
        const size_t k = order[i];
        ptr[row[k] + 1]++;
        c[i] = (int)col[k];
        v[i] = make_double2(values[2 * k], values[2 * k + 1]);
    }
    for (size_t r = 0; r < n_rows; r++) ptr[r + 1] += ptr[r];
    te_matrix* A = new te_matrix();
    A->ctx = ctx;
    A->n_rows = (int64_t)n_rows;
    A->n_cols = (int64_t)n_cols;
    A->nnz = (int64_t)nnz;
    A->norm1 = colVal.empty() ? 0.0 : *std::max_element(colVal.begin(), colVal.end());
    A->norm1_valid = true;
    int rc = matrix_upload(ctx, A, ptr, c, v);
    if (rc != TE_OK) {
        te_matrix_destroy(A);
        return rc;
    }
    *out = A;
    return TE_OK;
}

extern "C" void te_matrix_destroy(te_matrix* A) {
    if (!A) return;
    if (A->row_ptr) hipFree(A->row_ptr);
    if (A->col) hipFree(A->col);
    if (A->val) hipFree(A->val);
    if (A->diag_re) hipFree(A->diag_re);
    if (A->row_block) hipFree(A->row_block);
    free_parts(A);
    free_vd(A);
    delete A;
}

extern "C" int te_matrix_storage(const te_matrix* A, int* format, size_t* n_values, size_t* stream_bytes) {
    if (!A) return TE_ERR_ARG;
    const int f = A->use_vd ? (A->vd.sell ? 4 : A->vd.perm ? 3 : 2) : A->shape.kind;
    if (format) *format = f;
    if (n_values) *n_values = A->use_vd ? (size_t)A->vd.n_values : 0;
    if (stream_bytes) {
        const size_t n = (size_t)A->n_rows, ns = (size_t)A->vd.n_slices;
        *stream_bytes = f == 4   ? 4 * (size_t)A->vd.sell_entries + 4 * ((size_t)A->vd.sell_entries / te::kSlice) +
                                     16 * (ns + 1) + (A->vd.diag_re ? 8 : 16) * n +
                                     (A->vd.sell_ufix ? 8 * te::kSellFix * ns + 8 * (size_t)A->vd.sell_overflow : 0)
                        : f >= 2 ? 6 * (size_t)A->vd.nnz + 4 * (n + 1) + 16 * n
                                 : 20 * (size_t)A->nnz + 4 * (n + 1);
    }
    return TE_OK;
}

extern "C" int te_matrix_info(const te_matrix* A, size_t* n_rows, size_t* n_cols, size_t* nnz) {
    if (!A) return TE_ERR_ARG;
    if (n_rows) *n_rows = (size_t)A->n_rows;
    if (n_cols) *n_cols = (size_t)A->n_cols;
    if (nnz) *nnz = (size_t)A->nnz;
    return TE_OK;
}

namespace te {
__global__ void colsum_kernel(int64_t n_rows, const int* ptr, const int* col, const double2* val, double* out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    for (int k = ptr[r]; k < ptr[r + 1]; k++) atomicAdd(&out[col[k]], hypot(val[k].x, val[k].y));
}
}  // namespace te

extern "C" int te_matrix_norm1(te_matrix* A, double* out) {
    if (!A || !out) return TE_ERR_ARG;
    te_ctx* ctx = A->ctx;
    if (!A->norm1_valid) {
        // Column sums on the device (atomic, so the summation order is not the reference's
        // COO order; the result can differ from it in the last bit).
        hipSetDevice(ctx->device);
        double* cs = nullptr;
        TE_HIP(ctx, hipMalloc(&cs, sizeof(double) * A->n_cols));
        TE_HIP(ctx, hipMemsetAsync(cs, 0, sizeof(double) * A->n_cols, ctx->stream));
        const int64_t blocks = (A->n_rows + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, ctx->stream, A->n_rows, A->row_ptr,
                           A->col, A->val, cs);
        std::vector<double> h(A->n_cols);
        TE_HIP(ctx, hipMemcpyAsync(h.data(), cs, sizeof(double) * A->n_cols, hipMemcpyDeviceToHost, ctx->stream));
        TE_HIP(ctx, hipStreamSynchronize(ctx->stream));
        hipFree(cs);
        A->norm1 = h.empty() ? 0.0 : *std::max_element(h.begin(), h.end());
        A->norm1_valid = true;
    }
    *out = A->norm1;
    return TE_OK;
}

extern "C" int te_matrix_to_coo(te_matrix* A, double* values, size_t* col, size_t* row) {
    if (!A) return TE_ERR_ARG;
    te_ctx* ctx = A->ctx;
    hipSetDevice(ctx->device);
    std::vector<int> ptr(A->n_rows + 1), c(A->nnz);
    TE_HIP(ctx, hipMemcpy(ptr.data(), A->row_ptr, sizeof(int) * ptr.size(), hipMemcpyDeviceToHost));
    if (A->nnz) {
        TE_HIP(ctx, hipMemcpy(c.data(), A->col, sizeof(int) * A->nnz, hipMemcpyDeviceToHost));
        if (values) TE_HIP(ctx, hipMemcpy(values, A->val, sizeof(double2) * A->nnz, hipMemcpyDeviceToHost));
    }
    for (int64_t r = 0; r < A->n_rows; r++)
        for (int k = ptr[r]; k < ptr[r + 1]; k++) {
            if (col) col[k] = (size_t)c[k];
            if (row) row[k] = (size_t)r;
        }
    return TE_OK;
}

extern "C" int te_matrix_to_csr(te_matrix* A, double* values, int64_t* row_ptr, int* col) {
    if (!A) return TE_ERR_ARG;
    te_ctx* ctx = A->ctx;
    hipSetDevice(ctx->device);
    if (row_ptr) {
        std::vector<int> ptr(A->n_rows + 1);
        TE_HIP(ctx, hipMemcpy(ptr.data(), A->row_ptr, sizeof(int) * ptr.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < ptr.size(); i++) row_ptr[i] = ptr[i];
    }
    if (A->nnz && col) TE_HIP(ctx, hipMemcpy(col, A->col, sizeof(int) * A->nnz, hipMemcpyDeviceToHost));
    if (A->nnz && values) TE_HIP(ctx, hipMemcpy(values, A->val, sizeof(double2) * A->nnz, hipMemcpyDeviceToHost));
    return TE_OK;
}

extern "C" int te_spmv_device(te_matrix* A, double alpha_re, double alpha_im, const void* in_dev, void* out_dev) {
    if (!A || !in_dev || !out_dev) return TE_ERR_ARG;
    hipSetDevice(A->ctx->device);
    launch_spmv(A, make_double2(alpha_re, alpha_im), (const double2*)in_dev, (double2*)out_dev, A->ctx->stream);
    TE_HIP(A->ctx, hipGetLastError());
    return TE_OK;
}

extern "C" int te_spmv(te_matrix* A, double alpha_re, double alpha_im, const double* in, double* out) {
    if (!A || !in || !out) return TE_ERR_ARG;
    te_ctx* ctx = A->ctx;
    hipSetDevice(ctx->device);
    double2 *x = nullptr, *y = nullptr;
    TE_HIP(ctx, hipMalloc(&x, sizeof(double2) * A->n_cols));
    TE_HIP(ctx, hipMalloc(&y, sizeof(double2) * A->n_rows));
    TE_HIP(ctx, hipMemcpyAsync(x, in, sizeof(double2) * A->n_cols, hipMemcpyHostToDevice, ctx->stream));
    launch_spmv(A, make_double2(alpha_re, alpha_im), x, y, ctx->stream);
    TE_HIP(ctx, hipGetLastError());
    TE_HIP(ctx, hipMemcpyAsync(out, y, sizeof(double2) * A->n_rows, hipMemcpyDeviceToHost, ctx->stream));
    TE_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(x);
    hipFree(y);
    return TE_OK;
}

extern "C" int te_spmv_noalloc(te_matrix* A, double alpha_re, double alpha_im, const double* in, double* out) {
    using namespace te;
    if (!A || !in || !out) return TE_ERR_ARG;
    te_ctx* ctx = A->ctx;
    if (A->n_rows > kNoAllocRows || A->n_cols > kNoAllocRows)
        return fail(ctx, TE_ERR_UNSUPPORTED,
                    "te_spmv_noalloc: %lld x %lld matrix exceeds the %lld rows and columns of its staging; use "
                    "te_spmv", (long long)A->n_rows, (long long)A->n_cols, (long long)kNoAllocRows);
    double2* x = noalloc_scratch(ctx->device);
    if (!x) return fail(ctx, TE_ERR_HIP, "te_spmv_noalloc: no staging address on device %d", ctx->device);
    double2* y = x + kNoAllocRows;
    hipSetDevice(ctx->device);
    TE_HIP(ctx, hipMemcpyAsync(x, in, sizeof(double2) * A->n_cols, hipMemcpyHostToDevice, ctx->stream));
    launch_spmv(A, make_double2(alpha_re, alpha_im), x, y, ctx->stream);
    TE_HIP(ctx, hipGetLastError());
    TE_HIP(ctx, hipMemcpyAsync(out, y, sizeof(double2) * A->n_rows, hipMemcpyDeviceToHost, ctx->stream));
    TE_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TE_OK;
}

// =========================================================================================
// evolver
// =========================================================================================
struct te_evolver {
    te_ctx* ctx = nullptr;
    te_matrix* H = nullptr;
    te_evolve_params P{};
    int64_t n = 0;
    int m0 = 0;
    long long n_samples = 0;
    double matrixNorm = 0, numErrEst = 0, tolRate = 0;
    double2 *psi0 = nullptr, *U = nullptr, *sampled = nullptr, *tmp = nullptr, *tmp2 = nullptr;
    KrylovScalars* ks = nullptr;
    StepState* st = nullptr;
    StepState* st_host = nullptr;   // pinned
    double2 *wcoef = nullptr, *Y = nullptr;
    double* trace = nullptr;
    long long trace_cap = 0;
    std::vector<double> trace_host;
    double* obs_dev = nullptr;      // n_obs x n_samples
    std::vector<te_observable> obs;
    std::vector<int> diag_obs, generic_obs;
    double* diag_all = nullptr;     // |diag_obs| x n (freed once the byte form below exists)
    unsigned char* diag_idx = nullptr;  // n x idx_stride: row r's indices into diag_tab
    int idx_stride = 0;             // bytes per row: |diag_obs| rounded up to 16
    double* diag_tab = nullptr;     // |diag_obs| x kDiagTab distinct values
    double* diag_vals = nullptr;    // |diag_obs| x n_samples
    std::vector<double2*> vec_dev;  // per observable (vector kind)
    std::vector<double> host_vals;  // n_obs x n_samples, host-callback observables
    int stop_obs = -1;              // observable whose host callback requested the stop (last run)
    std::vector<size_t> recorded;   // per observable: samples recorded in the last run
    std::vector<double2> host_state;
    DevQuad quad;
    ReduceScratch rs;
    int grid_stream = 8;
    int grid_project = 8;
    std::vector<double2> e1;
    // Locality order of H's value-dictionary copy (null: the basis order). psi0, U, the
    // sampled states, tmp, diag_all and vec_dev are then in that order; tmp_ref and state_out
    // hold states mapped back to the basis order for callers.
    const int* perm = nullptr;      // new -> old
    const int* iperm = nullptr;     // old -> new
    double2* tmp_ref = nullptr;
    double2* state_out = nullptr;
    // optional kernel timing: HIP event pairs on the context stream, per kernel kind
    int prof_kinds = 0;              // bit k: time launches of kind k
    std::vector<hipEvent_t> prof_ev[4];
    size_t prof_used[4] = {0, 0, 0, 0};
    double prof_ms[4] = {0, 0, 0, 0};
    size_t prof_n[4] = {0, 0, 0, 0};
};

namespace te {

static void evolver_free(te_evolver* ev) {
    if (!ev) return;
    hipSetDevice(ev->ctx->device);
    for (auto& v : ev->prof_ev)
        for (hipEvent_t e : v) hipEventDestroy(e);
    double2* ptrs[] = {ev->psi0, ev->U, ev->sampled, ev->tmp, ev->tmp2, ev->wcoef, ev->Y, ev->tmp_ref, ev->state_out};
    for (double2* p : ptrs)
        if (p) hipFree(p);
    if (ev->ks) hipFree(ev->ks);
    if (ev->st) hipFree(ev->st);
    if (ev->st_host) hipHostFree(ev->st_host);
    if (ev->trace) hipFree(ev->trace);
    if (ev->obs_dev) hipFree(ev->obs_dev);
    if (ev->diag_all) hipFree(ev->diag_all);
    if (ev->diag_idx) hipFree(ev->diag_idx);
    if (ev->diag_tab) hipFree(ev->diag_tab);
    if (ev->diag_vals) hipFree(ev->diag_vals);
    free_quadrature(ev->quad);
    for (double2* p : ev->vec_dev)
        if (p) hipFree(p);
    reduce_scratch_free(ev->rs);
    delete ev;
}

constexpr size_t kProfPairs = 4096;

static void prof_begin(te_evolver* ev, int kind) {
    if (((ev->prof_kinds >> kind) & 1) && ev->prof_used[kind] + 2 <= ev->prof_ev[kind].size())
        hipEventRecord(ev->prof_ev[kind][ev->prof_used[kind]], ev->ctx->stream);
}
static void prof_end(te_evolver* ev, int kind) {
    if (((ev->prof_kinds >> kind) & 1) && ev->prof_used[kind] + 2 <= ev->prof_ev[kind].size()) {
        hipEventRecord(ev->prof_ev[kind][ev->prof_used[kind] + 1], ev->ctx->stream);
        ev->prof_used[kind] += 2;
    }
}
// After the stream is synchronised: sums the recorded pairs.
static void prof_collect(te_evolver* ev) {
    for (int k = 0; k < 4; k++) {
        ev->prof_ms[k] = 0.0;
        ev->prof_n[k] = ev->prof_used[k] / 2;
        for (size_t i = 0; i + 1 < ev->prof_used[k]; i += 2) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, ev->prof_ev[k][i], ev->prof_ev[k][i + 1]);
            ev->prof_ms[k] += ms;
        }
    }
}

template <class T>
static int dmalloc(te_ctx* ctx, T** p, size_t count) {
    TE_HIP(ctx, hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1)));
    return TE_OK;
}

#define TE_TRY(expr)                 \
    do {                             \
        int rc_ = (expr);            \
        if (rc_ != TE_OK) {          \
            evolver_free(ev);        \
            return rc_;              \
        }                            \
    } while (0)

// ---- diagonal observables as one-byte indices into tables of their distinct values ----
// The combine reads every diagonal observable once a row on every pass over the basis; most
// of them (occupation numbers, projectors, small-integer functions of the state) take few
// distinct values, so a byte per row and a table of at most kDiagTab values replace the 8-byte
// diagonal. The values the combine multiplies are the same doubles, bit for bit.
constexpr int kDiagSlots = 1024;                 // open-addressing slots per observable
constexpr unsigned long long kDiagEmpty = ~0ull;  // slot key of an empty slot

__device__ __forceinline__ unsigned diag_hash(unsigned long long k) {
    return (unsigned)((k * 0x9E3779B97F4A7C15ull) >> 54) & (kDiagSlots - 1);
}

// Inserts every value's bit pattern into its observable's slots; count[o] = distinct values,
// flag[0] = 1 when a table overflows (or a value has the empty slot's bit pattern). A slot is
// read before any atomic: after the first few rows every value is found without one.
__global__ void diag_insert_kernel(const double* __restrict__ d, int nd, int64_t n, unsigned long long* keys,
                                   int* count, int* flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)nd * n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int o = (int)(i / n);
        const unsigned long long k = (unsigned long long)__double_as_longlong(d[i]);
        if (k == kDiagEmpty) {
            *flag = 1;
            continue;
        }
        unsigned long long* T = keys + (size_t)o * kDiagSlots;
        unsigned h = diag_hash(k);
        for (int probe = 0;; probe++) {
            if (probe == kDiagSlots) {
                *flag = 1;
                break;
            }
            unsigned long long cur = __atomic_load_n(&T[h], __ATOMIC_RELAXED);
            if (cur == k) break;
            if (cur == kDiagEmpty) {
                cur = atomicCAS(&T[h], kDiagEmpty, k);
                if (cur == kDiagEmpty) {
                    if (atomicAdd(&count[o], 1) >= kDiagTab) *flag = 1;
                    break;
                }
                if (cur == k) break;
            }
            h = (h + 1) & (kDiagSlots - 1);
        }
    }
}

// idx[r * stride + o] = the table index of value d[o * n + r] (slot_index: the index of each
// occupied slot); the padding bytes of a row stay 0.
__global__ void diag_index_kernel(const double* __restrict__ d, int nd, int64_t n, const unsigned long long* keys,
                                  const unsigned char* __restrict__ slot_index, unsigned char* idx, int stride) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)nd * n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int o = (int)(i / n);
        const unsigned long long k = (unsigned long long)__double_as_longlong(d[i]);
        const unsigned long long* T = keys + (size_t)o * kDiagSlots;
        unsigned h = diag_hash(k);
        // present (diag_insert_kernel inserted every value without overflow); the probe count
        // is bounded all the same
        for (int probe = 0; probe < kDiagSlots && T[h] != k; probe++) h = (h + 1) & (kDiagSlots - 1);
        idx[(size_t)(i - (int64_t)o * n) * stride + o] = slot_index[(size_t)o * kDiagSlots + h];
    }
}

// Builds ev->diag_idx / ev->diag_tab from ev->diag_all when every diagonal observable has at
// most kDiagTab distinct values, and frees ev->diag_all; otherwise leaves the doubles in place.
static int compress_diagonals(te_evolver* ev) {
    te_ctx* ctx = ev->ctx;
    hipStream_t s = ctx->stream;
    const int nd = (int)ev->diag_obs.size();
    const int64_t n = ev->n;
    if (nd == 0 || !ev->diag_all) return TE_OK;
    const char* e = std::getenv("TE_DIAG_BYTES");                   // 0: keep the 8-byte diagonals
    if (e && e[0] == '0') return TE_OK;
    unsigned long long* keys = nullptr;
    int* cnt = nullptr;                  // nd counts, then the overflow flag
    unsigned char* slot_index = nullptr;
    TE_HIP(ctx, hipMalloc(&keys, sizeof(unsigned long long) * kDiagSlots * nd));
    auto done = [&](int r) {
        if (keys) hipFree(keys);
        if (cnt) hipFree(cnt);
        if (slot_index) hipFree(slot_index);
        return r;
    };
    if (hipMalloc(&cnt, sizeof(int) * (nd + 1)) != hipSuccess ||
        hipMalloc(&slot_index, (size_t)kDiagSlots * nd) != hipSuccess)
        return done(fail(ctx, TE_ERR_ALLOC, "diagonal observable tables"));
    if (hipMemsetAsync(keys, 0xff, sizeof(unsigned long long) * kDiagSlots * nd, s) != hipSuccess ||
        hipMemsetAsync(cnt, 0, sizeof(int) * (nd + 1), s) != hipSuccess)
        return done(fail(ctx, TE_ERR_HIP, "diagonal observable tables"));
    const unsigned grid = (unsigned)std::min<int64_t>(((int64_t)nd * n + 255) / 256, (int64_t)ctx->num_cus * 8);
    hipLaunchKernelGGL(diag_insert_kernel, dim3(grid), dim3(256), 0, s, ev->diag_all, nd, n, keys, cnt, cnt + nd);
    std::vector<unsigned long long> hk((size_t)kDiagSlots * nd);
    std::vector<int> hc(nd + 1);
    if (hipMemcpyAsync(hk.data(), keys, sizeof(unsigned long long) * hk.size(), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipMemcpyAsync(hc.data(), cnt, sizeof(int) * hc.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return done(fail(ctx, TE_ERR_HIP, "diagonal observable tables"));
    if (hc[nd] != 0) return done(TE_OK);                         // too many values: keep the doubles
    // table order: ascending bit patterns (deterministic)
    std::vector<unsigned char> hs((size_t)kDiagSlots * nd, 0);
    std::vector<double> tab((size_t)kDiagTab * nd, 0.0);
    for (int o = 0; o < nd; o++) {
        std::vector<std::pair<unsigned long long, int>> used;
        for (int h = 0; h < kDiagSlots; h++)
            if (hk[(size_t)o * kDiagSlots + h] != kDiagEmpty) used.push_back({hk[(size_t)o * kDiagSlots + h], h});
        std::sort(used.begin(), used.end());
        for (size_t q = 0; q < used.size(); q++) {
            hs[(size_t)o * kDiagSlots + used[q].second] = (unsigned char)q;
            std::memcpy(&tab[(size_t)o * kDiagTab + q], &used[q].first, sizeof(double));
        }
    }
    ev->idx_stride = (nd + 15) / 16 * 16;
    if (hipMalloc(&ev->diag_idx, (size_t)ev->idx_stride * n) != hipSuccess ||
        hipMalloc(&ev->diag_tab, sizeof(double) * kDiagTab * nd) != hipSuccess)
        return done(fail(ctx, TE_ERR_ALLOC, "diagonal observable tables"));
    if (hipMemcpyAsync(slot_index, hs.data(), hs.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(ev->diag_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, s) != hipSuccess)
        return done(fail(ctx, TE_ERR_HIP, "diagonal observable tables"));
    if (hipMemsetAsync(ev->diag_idx, 0, (size_t)ev->idx_stride * n, s) != hipSuccess)
        return done(fail(ctx, TE_ERR_HIP, "diagonal observable tables"));
    hipLaunchKernelGGL(diag_index_kernel, dim3(grid), dim3(256), 0, s, ev->diag_all, nd, n, keys, slot_index,
                       ev->diag_idx, ev->idx_stride);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return done(fail(ctx, TE_ERR_HIP, "diagonal observable tables"));
    hipFree(ev->diag_all);
    ev->diag_all = nullptr;
    return done(TE_OK);
}

static int evolver_create(te_ctx* ctx, te_matrix* H, const void* psi0, bool psi0_on_device,
                          const te_evolve_params* params, const te_observable* obs, int n_obs, te_evolver** out) {
    if (!ctx || !H || !psi0 || !params || !out || (n_obs > 0 && !obs) || n_obs < 0) return TE_ERR_ARG;
    *out = nullptr;
    hipSetDevice(ctx->device);
    const int64_t n = H->n_rows;
    if (n == 0) return fail(ctx, TE_ERR_FATAL, "Invalid Hilbertspace dimension");      // :39-43
    if (H->n_cols != n) return fail(ctx, TE_ERR_ARG, "Hamiltonian must be square");
    if (!(params->t > 0) || !(params->sampling_step > 0)) return fail(ctx, TE_ERR_ARG, "t and samplingStep must be > 0");
    if (params->m < 1) return fail(ctx, TE_ERR_ARG, "Krylov dimension must be >= 1");
    const int m = (int)std::min<int64_t>(params->m, n);                                  // :35
    if (m > kMaxKrylov) return fail(ctx, TE_ERR_UNSUPPORTED, "Krylov dimension %d exceeds %d", m, kMaxKrylov);

    te_evolver* ev = new te_evolver();
    ev->ctx = ctx;
    ev->H = H;
    ev->P = *params;
    ev->n = n;
    ev->m0 = m;
    double nrm1 = 0;
    TE_TRY(te_matrix_norm1(H, &nrm1));                                                    // :33
    ev->matrixNorm = nrm1;
    ev->n_samples = (long long)std::floor(params->t / params->sampling_step) + 1;          // :57
    ev->tolRate = params->tol / params->t;                                                // :287
    ev->numErrEst = (double)n * nrm1 * DBL_EPSILON;                                       // :289
    ev->grid_stream = combine_grid(ctx, n);
    ev->grid_project = project_grid(ctx, n);

    TE_TRY(dmalloc(ctx, &ev->psi0, n));
    TE_TRY(dmalloc(ctx, &ev->U, (size_t)n * (m + 1)));
    TE_TRY(dmalloc(ctx, &ev->sampled, n));
    TE_TRY(dmalloc(ctx, &ev->ks, 1));
    TE_TRY(dmalloc(ctx, &ev->st, 1));
    if (hipHostMalloc((void**)&ev->st_host, sizeof(StepState)) != hipSuccess) {
        evolver_free(ev);
        return fail(ctx, TE_ERR_ALLOC, "pinned allocation failed");
    }
    TE_TRY(dmalloc(ctx, &ev->wcoef, kMaxKrylov));
    TE_TRY(dmalloc(ctx, &ev->Y, (size_t)ev->n_samples * kMaxKrylov));
    ev->trace_cap = 1 << 20;
    TE_TRY(dmalloc(ctx, &ev->trace, ev->trace_cap));
    TE_TRY(dmalloc(ctx, &ev->obs_dev, (size_t)std::max(n_obs, 1) * ev->n_samples));
    TE_TRY(reduce_scratch_alloc(ctx, ev->rs,
                                std::max({ev->grid_stream, ev->grid_project, H->shape.grid, H->vd.grid,
                                          H->vd.sell_grid})));

    if (psi0_on_device)
        TE_TRY(hipMemcpy(ev->psi0, psi0, sizeof(double2) * n, hipMemcpyDeviceToDevice) == hipSuccess ? TE_OK
                                                                                                  : TE_ERR_HIP);
    else
        TE_TRY(hipMemcpy(ev->psi0, psi0, sizeof(double2) * n, hipMemcpyHostToDevice) == hipSuccess ? TE_OK
                                                                                                : TE_ERR_HIP);
    // |‖psi0‖ - 1| <= tol (:45-48)
    launch_nrm2sq(ev->psi0, n, ev->rs, 0, ctx->stream);
    double2 r;
    TE_TRY(hipMemcpyAsync(&r, ev->rs.result, sizeof(double2), hipMemcpyDeviceToHost, ctx->stream) == hipSuccess
               ? TE_OK
               : TE_ERR_HIP);
    TE_TRY(hipStreamSynchronize(ctx->stream) == hipSuccess ? TE_OK : TE_ERR_HIP);
    if (std::fabs(std::sqrt(r.x) - 1.0) > params->tol) {
        evolver_free(ev);
        return fail(ctx, TE_ERR_FATAL, "Initial vector is not normalized");
    }
    if (H->use_vd && H->vd.perm) {            // work in the order of H's value-dictionary copy
        ev->perm = H->vd.perm;
        ev->iperm = H->vd.iperm;
        TE_TRY(dmalloc(ctx, &ev->state_out, n));
        TE_TRY(hipMemcpyAsync(ev->sampled, ev->psi0, sizeof(double2) * n, hipMemcpyDeviceToDevice, ctx->stream) ==
                       hipSuccess
                   ? TE_OK
                   : TE_ERR_HIP);
        launch_gather(ev->sampled, ev->perm, ev->psi0, n, ctx->stream);
    }

    // observables
    ev->obs.assign(obs, obs + n_obs);
    ev->vec_dev.assign(n_obs, nullptr);
    ev->host_vals.assign((size_t)n_obs * ev->n_samples, 0.0);
    std::vector<te_matrix*> diag_m;
    for (int i = 0; i < n_obs; i++) {
        const te_observable& o = obs[i];
        if (o.kind == TE_OBS_SPARSE) {
            if (!o.matrix || o.matrix->n_rows != n || o.matrix->n_cols != n) {
                evolver_free(ev);
                return fail(ctx, TE_ERR_FATAL, "Incompatible dimensions");   // krylovObservables.cpp:182-185
            }
            if (o.matrix->diagonal) {
                ev->diag_obs.push_back(i);
                diag_m.push_back(o.matrix);
            } else {
                ev->generic_obs.push_back(i);
            }
        } else if (o.kind == TE_OBS_VECTOR) {
            if (!o.vector) {
                evolver_free(ev);
                return fail(ctx, TE_ERR_ARG, "null vector observable");
            }
            TE_TRY(dmalloc(ctx, &ev->vec_dev[i], n));
            double2* dst = ev->perm ? ev->sampled : ev->vec_dev[i];
            TE_TRY(hipMemcpyAsync(dst, o.vector, sizeof(double2) * n, hipMemcpyHostToDevice, ctx->stream) ==
                           hipSuccess
                       ? TE_OK
                       : TE_ERR_HIP);
            if (ev->perm) launch_gather(ev->sampled, ev->perm, ev->vec_dev[i], n, ctx->stream);
            TE_TRY(hipStreamSynchronize(ctx->stream) == hipSuccess ? TE_OK : TE_ERR_HIP);
            ev->generic_obs.push_back(i);
        } else if (o.kind == TE_OBS_HOST) {
            if (!o.callback) {
                evolver_free(ev);
                return fail(ctx, TE_ERR_ARG, "null observable callback");
            }
            ev->generic_obs.push_back(i);
        } else {
            evolver_free(ev);
            return fail(ctx, TE_ERR_ARG, "unknown observable kind %d", o.kind);
        }
    }
    if (!diag_m.empty()) {
        TE_TRY(dmalloc(ctx, &ev->diag_all, diag_m.size() * (size_t)n));
        TE_TRY(dmalloc(ctx, &ev->diag_vals, diag_m.size() * (size_t)ev->n_samples));
        for (size_t i = 0; i < diag_m.size(); i++) {
            if (ev->perm)
                launch_gather_re(diag_m[i]->diag_re, ev->perm, ev->diag_all + i * n, n, ctx->stream);
            else
                TE_TRY(hipMemcpyAsync(ev->diag_all + i * n, diag_m[i]->diag_re, sizeof(double) * n,
                                      hipMemcpyDeviceToDevice, ctx->stream) == hipSuccess
                           ? TE_OK
                           : TE_ERR_HIP);
        }
        TE_TRY(hipStreamSynchronize(ctx->stream) == hipSuccess ? TE_OK : TE_ERR_HIP);
        TE_TRY(compress_diagonals(ev));
    }
    if (!ev->generic_obs.empty()) {
        TE_TRY(dmalloc(ctx, &ev->tmp, n));
        TE_TRY(dmalloc(ctx, &ev->tmp2, n));
        if (ev->perm) TE_TRY(dmalloc(ctx, &ev->tmp_ref, n));
    }

    // quadrature tables
    TE_TRY(upload_quadrature(ctx, ev->quad));
    ev->e1.assign(kMaxKrylov, make_double2(0.0, 0.0));
    ev->e1[0] = make_double2(1.0, 0.0);
    ev->host_state.resize(ev->generic_obs.empty() ? 0 : n);
    *out = ev;
    return TE_OK;
}

// Evaluates the generic observables on the state in ev->tmp for sample `sample`.
// Returns TE_ERR_STOP when a host observable requests a stop.
static int eval_generic(te_evolver* ev, long long sample) {
    te_ctx* ctx = ev->ctx;
    hipStream_t s = ctx->stream;
    bool host_copied = false;
    // sparse observables and host callbacks see the state in the basis order
    const double2* ref = ev->tmp;
    if (ev->perm) {
        launch_gather(ev->tmp, ev->iperm, ev->tmp_ref, ev->n, s);
        ref = ev->tmp_ref;
    }
    for (int i : ev->generic_obs) {
        const te_observable& o = ev->obs[i];
        double* dst = ev->obs_dev + (size_t)i * ev->n_samples + sample;
        if (o.kind == TE_OBS_SPARSE) {
            launch_spmv(o.matrix, make_double2(1.0, 0.0), ref, ev->tmp2, s);
            launch_zdotc(ref, ev->tmp2, ev->n, ev->rs, 1, s);
            launch_store_obs(ev->rs.result, 1, 0, dst, s);
        } else if (o.kind == TE_OBS_VECTOR) {
            launch_zdotc(ev->tmp, ev->vec_dev[i], ev->n, ev->rs, 1, s);
            launch_store_obs(ev->rs.result, 1, 1, dst, s);
        } else {
            if (!host_copied) {
                TE_HIP(ctx, hipMemcpyAsync(ev->host_state.data(), ref, sizeof(double2) * ev->n,
                                           hipMemcpyDeviceToHost, s));
                TE_HIP(ctx, hipStreamSynchronize(s));
                host_copied = true;
            }
            double v = 0.0;
            if (o.callback(o.user, (const double*)ev->host_state.data(), (size_t)ev->n, &v) != 0) {
                ev->stop_obs = i;
                return TE_ERR_STOP;
            }
            ev->host_vals[(size_t)i * ev->n_samples + sample] = v;
        }
    }
    TE_HIP(ctx, hipGetLastError());
    return TE_OK;
}

static CombineArgs combine_base(te_evolver* ev, int m) {
    CombineArgs a{};
    a.n = ev->n;
    a.U = ev->U;
    a.ld = ev->n;
    a.m = m;
    a.wcoef = ev->wcoef;
    a.U0 = ev->U;
    a.Y = ev->Y;
    a.diag = ev->diag_all;
    a.obs_values = ev->diag_vals;
    a.n_samples = ev->n_samples;
    a.partials = ev->rs.partials;
    return a;
}

// Materialises sample `sample` into ev->tmp (U y_sample).
static void materialise(te_evolver* ev, int m, long long sample) {
    CombineArgs a = combine_base(ev, m);
    a.write_w = 0;
    a.y_base = sample;
    a.S = 1;
    a.write_last = 1;
    a.last_out = ev->tmp;
    a.nobs = 0;
    launch_combine(a, ev->grid_stream, ev->ctx->stream);
}

// Diagonal observables for samples [base, base+S) and, when requested, the new currentVec
// and the last sampled state. Launches go observable chunk by sample chunk; the final one
// covers the step's last sample and is the only one that writes U[:,0] and the state.
static void combine_step(te_evolver* ev, int m, long long base, int S, bool write_w, bool write_last) {
    const int nd = (int)ev->diag_obs.size();
    struct Launch {
        long long y0;
        int S, o0, no;
        bool tiled;
    };
    std::vector<Launch> L;
    for (int o0 = 0; o0 < nd; o0 += combine_obs_per_launch()) {
        const int no = std::min(nd - o0, combine_obs_per_launch());
        int sc = combine_samples_per_launch(no);
        // more samples than one pass of combine_kernel takes: the tiled pass, when it applies
        // (TE_COMBINE_TILED=0: passes of combine_kernel only; read at every step)
        const char* te = std::getenv("TE_COMBINE_TILED");
        const int tsc = (te && te[0] == '0') ? 0 : combine_tiled_samples(no);
        const bool tiled = S > sc && tsc > sc;
        if (tiled) sc = tsc;
        for (int s0 = 0; s0 < S; s0 += sc) L.push_back({base + s0, std::min(sc, S - s0), o0, no, tiled});
    }
    if (L.empty()) {
        if (!write_w && !(write_last && S > 0)) return;
        if (S > 0) L.push_back({base + S - 1, 1, 0, 0, false});
        else L.push_back({base, 0, 0, 0, false});
    }
    for (size_t i = 0; i < L.size(); i++) {
        const bool last = i + 1 == L.size();
        CombineArgs a = combine_base(ev, m);
        a.y_base = L[i].y0;
        a.S = L[i].S;
        a.nobs = L[i].no;
        a.obs_base = L[i].o0;
        a.diag = ev->diag_all ? ev->diag_all + (size_t)L[i].o0 * ev->n : nullptr;
        a.diag_idx = ev->diag_idx ? ev->diag_idx + L[i].o0 : nullptr;      // launches start on multiples of 16
        a.idx_stride = ev->idx_stride;
        a.diag_tab = ev->diag_tab ? ev->diag_tab + (size_t)L[i].o0 * kDiagTab : nullptr;
        a.write_w = (last && write_w) ? 1 : 0;
        a.write_last = (last && write_last && S > 0) ? 1 : 0;
        a.last_out = ev->sampled;
        if (L[i].tiled) launch_combine_tiled(a, ev->grid_stream, ev->ctx->stream);
        else launch_combine(a, ev->grid_stream, ev->ctx->stream);
    }
}

}  // namespace te

extern "C" int te_evolver_create(te_ctx* ctx, te_matrix* H, const double* psi0, const te_evolve_params* params,
                                 const te_observable* obs, int n_obs, te_evolver** out) {
    return te::evolver_create(ctx, H, psi0, false, params, obs, n_obs, out);
}

extern "C" int te_evolver_create_device(te_ctx* ctx, te_matrix* H, const void* psi0_dev,
                                        const te_evolve_params* params, const te_observable* obs, int n_obs,
                                        te_evolver** out) {
    return te::evolver_create(ctx, H, psi0_dev, true, params, obs, n_obs, out);
}

extern "C" void te_evolver_destroy(te_evolver* ev) { te::evolver_free(ev); }

extern "C" int te_evolver_num_samples(const te_evolver* ev, size_t* n_samples) {
    if (!ev || !n_samples) return TE_ERR_ARG;
    *n_samples = (size_t)ev->n_samples;
    return TE_OK;
}

extern "C" int te_evolver_state_device(te_evolver* ev, const void** state_dev) {
    if (!ev || !state_dev) return TE_ERR_ARG;
    *state_dev = ev->perm ? ev->state_out : ev->sampled;
    return TE_OK;
}

extern "C" int te_evolver_step_trace(te_evolver* ev, double* t_steps, size_t cap, size_t* n) {
    if (!ev || !n) return TE_ERR_ARG;
    *n = ev->trace_host.size();
    if (t_steps)
        for (size_t i = 0; i < std::min(cap, ev->trace_host.size()); i++) t_steps[i] = ev->trace_host[i];
    return TE_OK;
}

extern "C" int te_evolver_run(te_evolver* ev, double* obs_values, double* evolved_state, te_result* result) {
    using namespace te;
    if (!ev || !result) return TE_ERR_ARG;
    te_ctx* ctx = ev->ctx;
    hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const int64_t n = ev->n;
    const int n_obs = (int)ev->obs.size();
    const double2 alpha = make_double2(0.0, -1.0 * ev->P.exp_factor);                   // :31
    std::fill(ev->host_vals.begin(), ev->host_vals.end(), 0.0);
    ev->stop_obs = -1;
    ev->trace_host.clear();
    for (int k = 0; k < 4; k++) ev->prof_used[k] = 0;

    // ---- initial state (:59-64) and the sample at t = 0 (:305-313) ----
    StepState init{};
    init.t_step = 1.0 / 10.0;                                                           // :285
    init.m = ev->m0;
    init.index_samples = 1;
    *ev->st_host = init;
    KrylovScalars ks0{};
    ks0.s[0] = 1.0;
    TE_HIP(ctx, hipMemcpyAsync(ev->ks, &ks0, sizeof(ks0), hipMemcpyHostToDevice, s));
    TE_HIP(ctx, hipMemcpyAsync(ev->st, ev->st_host, sizeof(StepState), hipMemcpyHostToDevice, s));
    TE_HIP(ctx, hipMemcpyAsync(ev->U, ev->psi0, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
    TE_HIP(ctx, hipMemcpyAsync(ev->Y, ev->e1.data(), sizeof(double2) * kMaxKrylov, hipMemcpyHostToDevice, s));
    TE_HIP(ctx, hipMemsetAsync(ev->obs_dev, 0, sizeof(double) * std::max(n_obs, 1) * ev->n_samples, s));
    if (ev->diag_vals)
        TE_HIP(ctx, hipMemsetAsync(ev->diag_vals, 0, sizeof(double) * ev->diag_obs.size() * ev->n_samples, s));

    te_result R{};
    R.dim = (size_t)n;
    R.krylov_dim = (size_t)ev->m0;
    int statusCode = 0;
    bool stopped = false;
    if (!ev->generic_obs.empty()) {
        TE_HIP(ctx, hipMemcpyAsync(ev->tmp, ev->psi0, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
        int rc = eval_generic(ev, 0);
        if (rc == TE_ERR_STOP) {
            // status stays 0 and no sample is counted (:309-313); the observables listed before
            // the one that stopped have recorded sample 0, as sample() evaluates them in order (:120-126)
            stopped = true;
            R.num_samples = 0;
            combine_step(ev, 1, 0, 1, false, false);
            TE_HIP(ctx, hipMemcpyAsync(ev->sampled, ev->psi0, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
        } else if (rc != TE_OK) {
            return rc;
        }
    }
    if (!stopped) combine_step(ev, 1, 0, 1, false, true);
    TE_HIP(ctx, hipGetLastError());

    int m = ev->m0;
    LanczosArgs la{};
    la.n = n;
    la.row_block = ev->H->row_block;
    la.n_row_blocks = ev->H->n_row_blocks;
    la.ptr = ev->H->row_ptr;
    la.col = ev->H->col;
    la.val = ev->H->val;
    la.U = ev->U;
    la.ld = n;
    la.alpha = alpha;
    la.tolRate = ev->tolRate;
    la.ks = ev->ks;
    la.partials = ev->rs.partials;
    la.ticket = ev->rs.ticket;
    StepConsts C = step_consts(ev->P, ev->tolRate, ev->numErrEst, ev->n_samples, ev->quad);
    C.Y = ev->Y;
    C.wcoef = ev->wcoef;
    C.trace = ev->trace;
    C.trace_cap = ev->trace_cap;
    if (ev->P.fast_integration)
        log_msg(2, "Please note that a less accurate method for evaluating the error integral was used. We "
                   "recommend recomputing with accurate integration to establish the validity of the result.");

    StepState& h = *ev->st_host;
    int stalled = 0;
    while (!stopped && h.index_samples < ev->n_samples) {                            // :332
        la.m = m;
        for (int j = 0; j < m; j++) {                                               // :338
            la.j = j;
            prof_begin(ev, 0);
            launch_lanczos_apply(la, ev->H, s);
            prof_end(ev, 0);
            prof_begin(ev, 1);
            launch_lanczos_project(la, ev->grid_project, s);
            prof_end(ev, 1);
        }
        prof_begin(ev, 2);
        launch_step_control(C, ev->st, ev->ks, s);
        prof_end(ev, 2);
        TE_HIP(ctx, hipGetLastError());
        TE_HIP(ctx, hipMemcpyAsync(ev->st_host, ev->st, sizeof(StepState), hipMemcpyDeviceToHost, s));
        TE_HIP(ctx, hipStreamSynchronize(s));
        if (h.fatal == 1)
            return fail(ctx, TE_ERR_FATAL, "Error: No small enough time step found to meet tolerance requirements.");
        if (h.fatal == 2) return fail(ctx, TE_ERR_FATAL, "Internal error: eigen-solver did not converge");
        if (h.breakdown) log_msg(1, "***Lucky breakdown at Krylov dimension%d *** ", h.m);
        // A step that neither advances time nor takes a sample would repeat forever.
        stalled = (h.t_step == 0.0 && h.S == 0) ? stalled + 1 : 0;
        if (stalled >= 2) return fail(ctx, TE_ERR_FATAL, "time evolution makes no progress (t_step = 0)");
        m = h.m;
        // generic observables first: the last combine launch overwrites U[:,0]
        for (int si = 0; si < h.S && !ev->generic_obs.empty(); si++) {
            const long long sample = h.sample_base + si;
            materialise(ev, m, sample);
            const int rc = eval_generic(ev, sample);
            if (rc == TE_ERR_STOP) {                                                    // :425-436
                // the diagonal observables of the step's samples up to and including this one (those
                // listed after the stopping observable drop this sample below)
                combine_step(ev, m, h.sample_base, si + 1, false, false);
                TE_HIP(ctx, hipMemcpyAsync(ev->sampled, ev->tmp, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
                stopped = true;
                statusCode = 3;
                R.num_samples = (size_t)sample;
                R.evolved_time = h.t_now_prev;
                R.err = h.err_prev;
                R.n_steps = (size_t)h.n_steps;
                R.krylov_dim = (size_t)m;
                break;
            }
            if (rc != TE_OK) return rc;
        }
        if (stopped) break;
        prof_begin(ev, 3);
        combine_step(ev, m, h.sample_base, h.S, h.write_w != 0, true);
        prof_end(ev, 3);
        TE_HIP(ctx, hipGetLastError());
    }

    if (!stopped) {
        R.err = h.err;
        R.evolved_time = h.t_now;
        R.num_samples = (size_t)h.index_samples;
        R.n_steps = (size_t)h.n_steps;
        R.krylov_dim = (size_t)h.m;
        // norm check (:446-452) and status aggregation (:459-504)
        launch_nrm2sq(ev->sampled, n, ev->rs, 0, s);
        double2 r2;
        TE_HIP(ctx, hipMemcpyAsync(&r2, ev->rs.result, sizeof(double2), hipMemcpyDeviceToHost, s));
        TE_HIP(ctx, hipStreamSynchronize(s));
        statusCode = final_status(h, ev->numErrEst, ev->P.tol, std::sqrt(r2.x));
    }
    R.status_code = statusCode;

    // samples each observable recorded: all of the run's, and after a stop request the stop
    // sample as well for the observables listed before the one that requested it
    ev->recorded.assign((size_t)n_obs, R.num_samples);
    if (stopped && ev->stop_obs >= 0)
        for (int i = 0; i < ev->stop_obs; i++) ev->recorded[i] = R.num_samples + 1;
    // trace and outputs
    const size_t nt = (size_t)std::min<long long>(h.n_steps, ev->trace_cap);
    ev->trace_host.resize(nt);
    if (nt) TE_HIP(ctx, hipMemcpyAsync(ev->trace_host.data(), ev->trace, sizeof(double) * nt, hipMemcpyDeviceToHost, s));
    if (obs_values && n_obs > 0) {
        TE_HIP(ctx, hipMemcpyAsync(obs_values, ev->obs_dev, sizeof(double) * n_obs * ev->n_samples,
                                   hipMemcpyDeviceToHost, s));
        for (size_t k = 0; k < ev->diag_obs.size(); k++)
            TE_HIP(ctx, hipMemcpyAsync(obs_values + (size_t)ev->diag_obs[k] * ev->n_samples,
                                       ev->diag_vals + k * ev->n_samples, sizeof(double) * ev->n_samples,
                                       hipMemcpyDeviceToHost, s));
    }
    const double2* state = ev->sampled;
    if (ev->perm) {                                   // back to the basis order
        launch_gather(ev->sampled, ev->iperm, ev->state_out, n, s);
        state = ev->state_out;
    }
    if (evolved_state)
        TE_HIP(ctx, hipMemcpyAsync(evolved_state, state, sizeof(double2) * n, hipMemcpyDeviceToHost, s));
    TE_HIP(ctx, hipStreamSynchronize(s));
    prof_collect(ev);
    if (obs_values)
        for (int i : ev->generic_obs)
            if (ev->obs[i].kind == TE_OBS_HOST)
                std::memcpy(obs_values + (size_t)i * ev->n_samples, ev->host_vals.data() + (size_t)i * ev->n_samples,
                            sizeof(double) * ev->n_samples);
        // a sample an observable did not record reads 0
        for (int i = 0; i < n_obs; i++)
            for (size_t k = ev->recorded[i]; k < (size_t)ev->n_samples; k++) obs_values[(size_t)i * ev->n_samples + k] = 0.0;
    *result = R;
    return TE_OK;
}

extern "C" int te_evolver_diag_bytes(const te_evolver* ev, size_t* bytes_per_row) {
    if (!ev || !bytes_per_row) return TE_ERR_ARG;
    *bytes_per_row = ev->diag_idx ? (size_t)ev->idx_stride : sizeof(double) * ev->diag_obs.size();
    return TE_OK;
}

extern "C" int te_evolver_obs_recorded(const te_evolver* ev, size_t* counts) {
    if (!ev || !counts) return TE_ERR_ARG;
    for (size_t i = 0; i < ev->obs.size(); i++) counts[i] = i < ev->recorded.size() ? ev->recorded[i] : 0;
    return TE_OK;
}

extern "C" int te_evolver_set_profiling(te_evolver* ev, int kinds) {
    if (!ev || kinds < 0 || kinds > 15) return TE_ERR_ARG;
    ev->prof_kinds = kinds;
    if (kinds && ev->prof_ev[0].empty()) {
        hipSetDevice(ev->ctx->device);
        for (auto& v : ev->prof_ev) {
            v.resize(2 * te::kProfPairs);
            for (hipEvent_t& e : v) TE_HIP(ev->ctx, hipEventCreate(&e));
        }
    }
    return TE_OK;
}

extern "C" int te_evolver_kernel_time(const te_evolver* ev, int kind, double* total_ms, size_t* launches) {
    if (!ev || kind < 0 || kind > 6 || !total_ms || !launches) return TE_ERR_ARG;
    if (kind >= 4) {          // step-controller phases, from the device wall clock
        int khz = 0;
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ev->ctx->device);
        *total_ms = khz > 0 ? (double)ev->st_host->phase_ticks[kind - 4] / khz : 0.0;
        *launches = (size_t)ev->st_host->n_steps;
        return TE_OK;
    }
    *total_ms = ev->prof_ms[kind];
    *launches = ev->prof_n[kind];
    return TE_OK;
}
