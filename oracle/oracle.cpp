// ORACLE — test infrastructure only.
//
// A CPU restatement of TimeEvolver's Krylov time-evolution path, used by tests/ as the
// checker of the HIP product path and by bench.py as the timed CPU baseline ("port").
// Nothing in timeevolver_amd/ links, loads or calls this file.
//
// Each function cites the reference file:line it restates (paths relative to the
// reference repository root). The reference itself cannot be built here (it needs
// Boost, mkl.h, cblas.h and lapacke.h, none of which are installed), so this is an
// independent restatement, pinned by tests/test_oracle_*.py against:
//   * the reference's golden observables (output/*.h5, output/*.csv),
//   * closed forms (two-oscillator Rabi formula, free-boson ring),
//   * dense linear algebra (numpy eigh, scipy expm) and mpmath quadrature,
//   * hand traces and closed-form counts of the basis and the CSR matrix.
// Parity unpinned (no reference artefact fixes them): the exact Krylov step
// schedule, n_steps and err to all digits (they depend on Boost's tanh_sinh
// internals); these are graded only through state/observable parity within tol.
//
// Compile with -ffp-contract=off: the matrix builder must reproduce the reference's
// separately rounded products bit for bit.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <unordered_map>
#include <vector>
#include <dlfcn.h>
#include <omp.h>

using cd = std::complex<double>;

namespace {

// ----------------------------------------------------------------------------------
// MKL backend (CPU baseline only). The reference is built against MKL with 64-bit
// integers (CMakeLists.txt:52 Intel10_64ilp; core/mathHeader.h:6-7 MKL_INT = size_t):
// its spMV is mkl_sparse_z_mv (core/matrixDataTypes.cpp:169-174) on a handle made by
// create_coo -> convert_csr -> order -> mv/memory hints -> optimize (:196-224), and the
// Lanczos and state updates are cblas calls (core/krylovTimeEvolver.cpp:382, :421,
// :523-571; core/krylovObservables.cpp:187-188). No MKL header is installed here, so the
// runtime library (libmkl_rt) is opened with dlopen and the documented C interfaces are
// declared below; the ILP64 interface and the GNU OpenMP threading layer are selected
// before any other call. Enabled only by or_mkl_enable; every evolver entry then routes
// its products and level-1/2 kernels through MKL.
// ----------------------------------------------------------------------------------
namespace mkl {
using i64 = long long;
struct Z { double re, im; };
struct Descr { int type, mode, diag; };
enum { INDEX_BASE_ZERO = 0, OP_NON_TRANSPOSE = 10, MATRIX_GENERAL = 20, FILL_FULL = 42, DIAG_NON_UNIT = 50,
       MEMORY_AGGRESSIVE = 81, INTERFACE_ILP64 = 1, THREADING_GNU = 3, COL_MAJOR = 102, NO_TRANS = 111 };
struct Api {
    void* so = nullptr;
    int (*set_interface)(int) = nullptr;
    int (*set_threading)(int) = nullptr;
    void (*set_threads)(int) = nullptr;
    void (*version)(char*, int) = nullptr;
    int (*create_coo)(void**, int, i64, i64, i64, i64*, i64*, Z*) = nullptr;
    int (*convert_csr)(void*, int, void**) = nullptr;
    int (*order)(void*) = nullptr;
    int (*mv_hint)(void*, int, Descr, i64) = nullptr;
    int (*memory_hint)(void*, int) = nullptr;
    int (*optimize)(void*) = nullptr;
    int (*mv)(int, Z, void*, Descr, const Z*, Z, Z*) = nullptr;
    int (*destroy)(void*) = nullptr;
    void (*zaxpy)(i64, const void*, const void*, i64, void*, i64) = nullptr;
    void (*zdotc_sub)(i64, const void*, i64, const void*, i64, void*) = nullptr;
    double (*dznrm2)(i64, const void*, i64) = nullptr;
    void (*zscal)(i64, const void*, void*, i64) = nullptr;
    void (*zcopy)(i64, const void*, i64, void*, i64) = nullptr;
    void (*zgemv)(int, int, i64, i64, const void*, const void*, i64, const void*, i64, const void*, void*,
                  i64) = nullptr;
    char ver[256] = {0};
};

Api g_api;
const Api* g_mkl = nullptr;              // non-null: MKL routes are on

template <class F>
bool sym(F& f, const char* name) {
    f = reinterpret_cast<F>(dlsym(g_api.so, name));
    return f != nullptr;
}

// Opens libmkl_rt (path given, or the loader's search path) and resolves every entry.
int enable(const char* path, int threads) {
    if (g_api.ver[0]) {                   // loaded before: switch back on
        g_api.set_threads(threads > 0 ? threads : omp_get_max_threads());
        g_mkl = &g_api;
        return 0;
    }
    const char* cands[] = {path, "libmkl_rt.so.1", "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1"};
    for (const char* c : cands)
        if (c && c[0] && (g_api.so = dlopen(c, RTLD_NOW | RTLD_GLOBAL))) break;
    if (!g_api.so) return 1;
    bool ok = sym(g_api.set_interface, "MKL_Set_Interface_Layer") && sym(g_api.set_threading, "MKL_Set_Threading_Layer") &&
              sym(g_api.set_threads, "MKL_Set_Num_Threads") && sym(g_api.version, "MKL_Get_Version_String") &&
              sym(g_api.create_coo, "mkl_sparse_z_create_coo") && sym(g_api.convert_csr, "mkl_sparse_convert_csr") &&
              sym(g_api.order, "mkl_sparse_order") && sym(g_api.mv_hint, "mkl_sparse_set_mv_hint") &&
              sym(g_api.memory_hint, "mkl_sparse_set_memory_hint") && sym(g_api.optimize, "mkl_sparse_optimize") &&
              sym(g_api.mv, "mkl_sparse_z_mv") && sym(g_api.destroy, "mkl_sparse_destroy") &&
              sym(g_api.zaxpy, "cblas_zaxpy") && sym(g_api.zdotc_sub, "cblas_zdotc_sub") &&
              sym(g_api.dznrm2, "cblas_dznrm2") && sym(g_api.zscal, "cblas_zscal") && sym(g_api.zcopy, "cblas_zcopy") &&
              sym(g_api.zgemv, "cblas_zgemv");
    if (!ok) return 2;
    if (g_api.set_interface(INTERFACE_ILP64) != INTERFACE_ILP64) return 3;
    g_api.set_threading(THREADING_GNU);
    g_api.set_threads(threads > 0 ? threads : omp_get_max_threads());
    g_api.version(g_api.ver, (int)sizeof(g_api.ver) - 1);
    g_mkl = &g_api;
    return 0;
}

Descr general() { return Descr{MATRIX_GENERAL, FILL_FULL, DIAG_NON_UNIT}; }
}  // namespace mkl

// ----------------------------------------------------------------------------------
// Basis: helper/Basis.cpp:37-120 (basis), :131-182 (tensorBasis), :13-17 (hash table)
// ----------------------------------------------------------------------------------
struct VecHash {
    size_t operator()(const std::vector<int>& v) const {
        uint64_t h = 1469598103934665603ull;
        for (int x : v) { h ^= (uint32_t)x; h *= 1099511628211ull; }
        return (size_t)h;
    }
};

struct Basis {
    int K = 0;
    long D = 0;
    std::vector<int> occ;                    // D x K, row-major
    bool indexed = false;
    std::unordered_map<std::vector<int>, long, VecHash> table;

    void build_table() {
        table.clear();
        table.reserve((size_t)D * 2 + 16);
        for (long i = 0; i < D; i++) {
            std::vector<int> v(occ.begin() + i * K, occ.begin() + (i + 1) * K);
            table.emplace(std::move(v), i);   // insert keeps the first index (Basis.cpp:15-16)
        }
        indexed = true;
    }
    long find(const int* e) const {
        std::vector<int> v(e, e + K);
        auto it = table.find(v);
        return it == table.end() ? -1 : it->second;
    }
};

// helper/Basis.cpp:37-120. The list starts as [zero vector]; for k = K..2 every state of
// the current list spawns, in list order, copies with e[k-1] = quant, 2*quant, ...,
// upper, which are appended behind the whole current list (writeIndex is never reset,
// :57, :84-85). upper = min(N - |s|, capacity) for the last nbQudits modes (:69-76).
// Finally e[0] = N - |s|, dropping states whose e[0] exceeds capacity when every mode is
// a qudit mode (:98-107).
Basis* make_basis(int N, int K, int nbQudits, int capacity, int quant) {
    Basis* b = new Basis();
    b->K = K;
    std::vector<int> cur(K, 0);          // flat list of states, K ints each
    std::vector<int> sums(1, 0);
    for (int k = K; k > 1; k--) {
        size_t old_len = sums.size();
        for (size_t l = 0; l < old_len; l++) {
            int occN = sums[l];
            int upper = (k > K - nbQudits) ? std::min(N - occN, capacity) : N - occN;
            for (int o = quant; o <= upper; o += quant) {
                size_t base = cur.size();
                cur.insert(cur.end(), cur.begin() + l * K, cur.begin() + (l + 1) * K);
                cur[base + (k - 1)] = o;
                sums.push_back(occN + o);
            }
        }
    }
    size_t n = sums.size();
    b->occ.reserve(n * K);
    for (size_t l = 0; l < n; l++) {
        int missing = N - sums[l];
        if (K != nbQudits || missing <= capacity) {
            size_t base = b->occ.size();
            b->occ.insert(b->occ.end(), cur.begin() + l * K, cur.begin() + (l + 1) * K);
            b->occ[base] = missing;
        }
    }
    b->D = (long)(b->occ.size() / (size_t)std::max(K, 1));
    if (K == 0) b->D = 0;
    return b;
}

// helper/Basis.cpp:131-182: element i*len2 + j = concat(b1[i], b2[j]) with
// b1 = basis(N1,K1,0,0) and b2 = basis(N2,K2,K2,C2). When len2 == 0 the reference
// reports numberElements = 0 (:143-153).
Basis* make_tensor_basis(int N1, int K1, int N2, int K2, int C2) {
    Basis* b1 = make_basis(N1, K1, 0, 0, 1);
    Basis* b2 = make_basis(N2, K2, K2, C2, 1);
    Basis* b = new Basis();
    b->K = K1 + K2;
    long l1 = b1->D, l2 = b2->D;
    b->D = l1 * l2;
    b->occ.resize((size_t)b->D * b->K);
    for (long i = 0; i < l1; i++)
        for (long j = 0; j < l2; j++) {
            int* dst = &b->occ[(size_t)(i * l2 + j) * b->K];
            for (int a = 0; a < K1; a++) dst[a] = b1->occ[i * K1 + a];
            for (int a = 0; a < K2; a++) dst[K1 + a] = b2->occ[j * K2 + a];
        }
    delete b1;
    delete b2;
    return b;
}

// ----------------------------------------------------------------------------------
// Operator -> sparse matrix: helper/hamiltonian.cpp:318-401 (createMatrix),
// :410-434 (creationOperator), :441-456 (annihilationOperator), :308-311 (scalar mult)
// ----------------------------------------------------------------------------------
struct Term {
    cd coef;
    std::vector<int> ops, modes, caps;
};

struct Coo {
    long n = 0;
    std::vector<cd> val;
    std::vector<long> col, row;
};

// Applies one opTerm to basis state e (K entries); returns the coefficient and leaves the
// image state in e. Mirrors the reference's multiplication sequence exactly:
// coef starts at (1,0), each operator multiplies it by sqrt(n+1.0) or sqrt(n) (a real
// factor: both components scaled), then the term coefficient multiplies it as a complex
// number (hamiltonian.cpp:343-358).
static cd apply_term(const Term& t, int* e) {
    cd coef(1.0, 0.0);
    for (size_t i = 0; i < t.ops.size(); i++) {
        int mode = t.modes[i];
        int n = e[mode];
        if (t.ops[i] == 1) {                       // creation (:410-434)
            int c = t.caps[i];
            if (c == 0) {
                coef = coef * std::sqrt(n + 1.0);
                e[mode] = n + 1;
            } else if (n + 1 > c) {
                coef = 0;
            } else {
                coef = coef * std::sqrt(n + 1.0);
                e[mode] = n + 1;
            }
        } else if (t.ops[i] == -1) {               // annihilation (:441-456)
            if (n == 0) coef = 0;
            else {
                coef = coef * std::sqrt((double)n);
                e[mode] = n - 1;
            }
        }
    }
    coef = coef * t.coef;                          // scalarMultiplication (:308-311)
    return coef;
}

// One row of createMatrix (hamiltonian.cpp:334-391): terms are applied in list order,
// found images are accumulated into a std::map keyed by column (first insertion, then
// +=), and the row is emitted in ascending column order skipping |v| < DBL_EPSILON.
template <class Finder>
static void build_row(const std::vector<Term>& terms, const int* state, int K, Finder find,
                      std::vector<std::pair<long, cd>>& out) {
    std::map<long, cd> perRow;
    std::vector<int> e(K);
    for (const Term& t : terms) {
        std::copy(state, state + K, e.begin());
        cd coef = apply_term(t, e.data());
        long idx = find(e.data());
        if (idx < 0) continue;
        auto it = perRow.find(idx);
        if (it == perRow.end()) perRow.insert(std::make_pair(idx, coef));
        else it->second += coef;
    }
    out.clear();
    for (auto& kv : perRow) {
        if (std::abs(kv.second) < std::numeric_limits<double>::epsilon()) continue;
        out.push_back(kv);
    }
}

Coo* build_operator(const Basis* b, const std::vector<Term>& terms) {
    Coo* m = new Coo();
    m->n = b->D;
    int nth = omp_get_max_threads();
    std::vector<std::vector<std::pair<long, cd>>> rows(b->D);
    auto finder = [b](const int* e) { return b->find(e); };
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nth)
    for (long j = 0; j < b->D; j++) {
        build_row(terms, &b->occ[(size_t)j * b->K], b->K, finder, rows[j]);
    }
    size_t nnz = 0;
    for (auto& r : rows) nnz += r.size();
    m->val.reserve(nnz); m->col.reserve(nnz); m->row.reserve(nnz);
    for (long j = 0; j < b->D; j++)
        for (auto& kv : rows[j]) {
            m->row.push_back(j);
            m->col.push_back(kv.first);
            m->val.push_back(kv.second);
        }
    return m;
}

// ----------------------------------------------------------------------------------
// CSR and the sparse backend seam: core/matrixDataTypes.cpp:169-190 (spMV),
// :67-79 (norm1). The COO arrays are converted to CSR ordered by (row, col), as
// mkl_sparse_convert_csr + mkl_sparse_order do (:213-224).
// ----------------------------------------------------------------------------------
struct Csr {
    long n = 0, ncols = 0;
    std::vector<long> ptr, col;
    std::vector<cd> val;
    void* mkl = nullptr;        // MKL handle (MKL backend only)
    Csr() = default;
    Csr(const Csr&) = delete;
    Csr& operator=(const Csr&) = delete;
    Csr(Csr&& o) noexcept : n(o.n), ncols(o.ncols), ptr(std::move(o.ptr)), col(std::move(o.col)),
                            val(std::move(o.val)), mkl(o.mkl) { o.mkl = nullptr; }
    Csr& operator=(Csr&& o) noexcept {
        std::swap(n, o.n); std::swap(ncols, o.ncols); ptr.swap(o.ptr); col.swap(o.col); val.swap(o.val);
        std::swap(mkl, o.mkl);
        return *this;
    }
    ~Csr() { if (mkl && mkl::g_api.destroy) mkl::g_api.destroy(mkl); }
};

// The reference's smatrix::initialize (core/matrixDataTypes.cpp:196-224) on the MKL backend:
// COO handle with 64-bit indices, converted to CSR, ordered, mv hint of 20000 calls,
// aggressive memory hint, optimize. The COO arrays only need to live until the conversion.
static int mkl_attach(Csr& A) {
    using namespace mkl;
    if (!g_mkl || A.mkl) return 0;
    const long nnz = A.ptr[A.n];
    if (nnz == 0) return 0;
    std::vector<i64> row(nnz), col(nnz);
#pragma omp parallel for schedule(static)
    for (long r = 0; r < A.n; r++)
        for (long k = A.ptr[r]; k < A.ptr[r + 1]; k++) {
            row[k] = r;
            col[k] = A.col[k];
        }
    void* coo = nullptr;
    if (g_mkl->create_coo(&coo, INDEX_BASE_ZERO, A.n, A.ncols, nnz, row.data(), col.data(),
                          reinterpret_cast<Z*>(A.val.data())) != 0)
        return 1;
    void* csr = nullptr;
    int rc = g_mkl->convert_csr(coo, OP_NON_TRANSPOSE, &csr);
    g_mkl->destroy(coo);
    if (rc != 0) return 2;
    g_mkl->order(csr);
    g_mkl->mv_hint(csr, OP_NON_TRANSPOSE, general(), 20000);
    g_mkl->memory_hint(csr, MEMORY_AGGRESSIVE);
    if (g_mkl->optimize(csr) != 0) {
        g_mkl->destroy(csr);
        return 3;
    }
    A.mkl = csr;
    return 0;
}

static Csr coo_to_csr(long n, long ncols, long nnz, const double* v, const long* col, const long* row) {
    Csr A;
    A.n = n; A.ncols = ncols;
    std::vector<long> order(nnz);
    for (long i = 0; i < nnz; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](long a, long b) {
        return row[a] != row[b] ? row[a] < row[b] : col[a] < col[b];
    });
    A.ptr.assign(n + 1, 0);
    A.col.resize(nnz); A.val.resize(nnz);
    for (long i = 0; i < nnz; i++) {
        long k = order[i];
        A.ptr[row[k] + 1]++;
        A.col[i] = col[k];
        A.val[i] = cd(v[2 * k], v[2 * k + 1]);
    }
    for (long r = 0; r < n; r++) A.ptr[r + 1] += A.ptr[r];
    return A;
}

// y = alpha * A x (matrixDataTypes.cpp:172-173: mkl_sparse_z_mv with beta = 0).
static void spmv(const Csr& A, cd alpha, const cd* x, cd* y) {
    if (A.mkl) {
        const mkl::Z a{alpha.real(), alpha.imag()}, zero{0.0, 0.0};
        mkl::g_mkl->mv(mkl::OP_NON_TRANSPOSE, a, A.mkl, mkl::general(), reinterpret_cast<const mkl::Z*>(x), zero,
                       reinterpret_cast<mkl::Z*>(y));
        return;
    }
#pragma omp parallel for schedule(static)
    for (long r = 0; r < A.n; r++) {
        cd s(0.0, 0.0);
        for (long k = A.ptr[r]; k < A.ptr[r + 1]; k++) s += A.val[k] * x[A.col[k]];
        y[r] = alpha * s;
    }
}

// Max absolute column sum, accumulated in COO order (matrixDataTypes.cpp:67-79).
static double norm1_coo(long ncols, long nnz, const double* v, const long* col) {
    std::vector<double> colVal(ncols, 0.0);
    for (long i = 0; i < nnz; i++) colVal[col[i]] += std::abs(cd(v[2 * i], v[2 * i + 1]));
    return colVal.empty() ? 0.0 : *std::max_element(colVal.begin(), colVal.end());
}

// ---- dense level-1 kernels with a fixed reduction order (per-thread blocks, combined
// in thread order), so results are deterministic for a given thread count ----
static cd zdotc(long n, const cd* x, const cd* y) {         // sum conj(x) y
    if (mkl::g_mkl) {
        cd r;
        mkl::g_mkl->zdotc_sub(n, x, 1, y, 1, &r);
        return r;
    }
    int nt = omp_get_max_threads();
    std::vector<cd> part(nt, cd(0.0, 0.0));
#pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num(), T = omp_get_num_threads();
        long lo = n * t / T, hi = n * (t + 1) / T;
        cd s(0.0, 0.0);
        for (long i = lo; i < hi; i++) s += std::conj(x[i]) * y[i];
        part[t] = s;
    }
    cd s(0.0, 0.0);
    for (auto& p : part) s += p;
    return s;
}
static double dznrm2(long n, const cd* x) {
    if (mkl::g_mkl) return mkl::g_mkl->dznrm2(n, x, 1);
    int nt = omp_get_max_threads();
    std::vector<double> part(nt, 0.0);
#pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num(), T = omp_get_num_threads();
        long lo = n * t / T, hi = n * (t + 1) / T;
        double s = 0.0;
        for (long i = lo; i < hi; i++) s += std::norm(x[i]);
        part[t] = s;
    }
    double s = 0.0;
    for (double p : part) s += p;
    return std::sqrt(s);
}
static void zaxpy(long n, cd a, const cd* x, cd* y) {
    if (mkl::g_mkl) {
        mkl::g_mkl->zaxpy(n, &a, x, 1, y, 1);
        return;
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) y[i] += a * x[i];
}
// y = V w, V column-major n x m (the reference's cblas_zgemv NoTrans, :382, :421).
static void gemv_tall(long n, int m, const cd* V, const cd* w, cd* y) {
    if (mkl::g_mkl) {
        const cd one(1.0, 0.0), zero(0.0, 0.0);
        mkl::g_mkl->zgemv(mkl::COL_MAJOR, mkl::NO_TRANS, n, m, &one, V, n, w, 1, &zero, y, 1);
        return;
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        cd s(0.0, 0.0);
        for (int j = 0; j < m; j++) s += V[(size_t)j * n + i] * w[j];
        y[i] = s;
    }
}

// ----------------------------------------------------------------------------------
// Krylov-space exponential. The reference Schur-factorises the m x m Hessenberg matrix
// with LAPACKE_zhseqr (matrixDataTypes.h:135-156, krylovTimeEvolver.cpp:365) and forms
// exp(tH)e1 = Z diag(e^{t lambda}) Z^H e1 (:685-702). The Lanczos matrix is
// anti-Hermitian tridiagonal: H[j][j] = a_j (purely imaginary up to rounding),
// H[j+1][j] = b_j, H[j][j+1] = -b_j (:542, :559-560). With D = diag((-i)^j),
// D H D^{-1} = -i T where T is real symmetric tridiagonal with T[j][j] = -Im(a_j) and
// off-diagonal b_j. Hence exp(tH) e1 = D^{-1} Q exp(-i t Lambda) Q^T e1 where
// T = Q Lambda Q^T. For a normal matrix the Schur form is this eigendecomposition,
// so the result equals the reference's up to rounding.
// ----------------------------------------------------------------------------------

// Implicit QL with Wilkinson shifts for a symmetric tridiagonal matrix (the textbook
// QL algorithm, e.g. Golub & Van Loan section 8.3). d: diagonal (n), e: off-diagonal
// (e[i] couples i and i+1, n-1 used), Q: n x n row-major, set to the eigenvectors
// (column k belongs to d[k]). Returns 0 on success, 1 if an eigenvalue did not converge.
static int tridiag_ql(int n, double* d, double* e, double* Q) {
    for (int i = 0; i < n * n; i++) Q[i] = 0.0;
    for (int i = 0; i < n; i++) Q[i * n + i] = 1.0;
    if (n == 1) return 0;
    e[n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        int iter = 0;
        while (true) {
            int mm;
            for (mm = l; mm < n - 1; mm++) {
                double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
                if (std::fabs(e[mm]) <= DBL_EPSILON * dd) break;
            }
            if (mm == l) break;
            if (++iter > 60) return 1;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            int i;
            bool early = false;
            for (i = mm - 1; i >= l; i--) {
                double f = s * e[i];
                double bb = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[mm] = 0.0;
                    early = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * bb;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - bb;
                for (int k = 0; k < n; k++) {
                    double qf = Q[k * n + i + 1];
                    Q[k * n + i + 1] = s * Q[k * n + i] + c * qf;
                    Q[k * n + i] = c * Q[k * n + i] - s * qf;
                }
            }
            if (early) continue;
            d[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
    return 0;
}

struct KrylovSpace {
    int m = 0;
    std::vector<double> lam;    // eigenvalues of T
    std::vector<double> Q;      // m x m row-major eigenvectors
    std::vector<double> g;      // Q[0][k]
    std::vector<double> p;      // Q[m-1][k] * Q[0][k]

    int factor(int mm, const cd* a, const double* b) {
        m = mm;
        lam.assign(m, 0.0);
        std::vector<double> e(m, 0.0);
        for (int j = 0; j < m; j++) lam[j] = -a[j].imag();
        for (int j = 0; j + 1 < m; j++) e[j] = b[j];
        Q.assign((size_t)m * m, 0.0);
        int info = tridiag_ql(m, lam.data(), e.data(), Q.data());
        g.resize(m); p.resize(m);
        for (int k = 0; k < m; k++) {
            g[k] = Q[k];
            p[k] = Q[(size_t)(m - 1) * m + k] * Q[k];
        }
        return info;
    }
    // |[exp(x H) e1]_{m-1}| = |sum_k p_k e^{-i x lam_k}| (the integrand core, :591)
    double last_abs(double x) const {
        double re = 0.0, im = 0.0;
        for (int k = 0; k < m; k++) {
            double s, c;
            sincos(x * lam[k], &s, &c);
            re += p[k] * c;
            im -= p[k] * s;
        }
        return std::hypot(re, im);
    }
    // out_j = [exp(x H) e1]_j = i^j sum_k Q[j][k] g_k e^{-i x lam_k}  (:685-702)
    void exp_e1(double x, cd* out) const {
        std::vector<cd> ph(m);
        for (int k = 0; k < m; k++) {
            double s, c;
            sincos(x * lam[k], &s, &c);
            ph[k] = cd(g[k] * c, -g[k] * s);
        }
        for (int j = 0; j < m; j++) {
            cd s(0.0, 0.0);
            for (int k = 0; k < m; k++) s += Q[(size_t)j * m + k] * ph[k];
            switch (j & 3) {                 // multiply by i^j exactly
                case 0: out[j] = s; break;
                case 1: out[j] = cd(-s.imag(), s.real()); break;
                case 2: out[j] = cd(-s.real(), -s.imag()); break;
                default: out[j] = cd(s.imag(), -s.real()); break;
            }
        }
    }
};

// ----------------------------------------------------------------------------------
// Quadrature for the error integral (krylovTimeEvolver.cpp:584-618).
// The reference uses Boost.Math (>= 1.75; CI pins 1.83.0): tanh_sinh<double> with
// integrate(f, a, b, termination = 1e-3, &error, &L1) and gauss<double,15>/<double,7>.
// Boost is not installed, so the published algorithms are restated:
//  * tanh-sinh (Takahasi & Mori 1974): x = c + d*tanh(pi/2 sinh t),
//    w = (pi/2) cosh t / cosh^2(pi/2 sinh t), trapezoidal sums with h = 1 on level 0,
//    each further level halving h and adding the odd nodes; error = |I_k - I_{k-1}|;
//    stop when error <= tol * L1 or after 14 refinements (Boost's default of 15 rows);
//    nodes whose abscissa complement falls below 4*DBL_MIN are never used, nor points
//    that round onto an end point of [a, b].
//  * Gauss-Legendre with 7 and 15 nodes (roots of P_n by Newton's method).
// ----------------------------------------------------------------------------------
struct TSNode { double xc, w; };   // abscissa complement 1 - tanh(u) and weight
struct TanhSinh {
    std::vector<std::vector<TSNode>> rows;  // row 0: t = 0,1,2,...; row k: odd multiples of 2^-k
    int max_levels = 15;
    TanhSinh() {
        const double half_pi = 1.5707963267948966;
        const double min_c = 4.0 * DBL_MIN;
        rows.resize(max_levels);
        auto node = [&](double t) {
            double u = half_pi * std::sinh(t);
            double ch = std::cosh(u);
            TSNode nd;
            nd.xc = 1.0 / (std::exp(u) * ch);
            nd.w = half_pi * std::cosh(t) / (ch * ch);
            return nd;
        };
        for (int t = 0;; t++) {
            TSNode nd = node((double)t);
            if (!(nd.xc >= min_c)) break;
            rows[0].push_back(nd);
        }
        for (int k = 1; k < max_levels; k++) {
            double h = std::ldexp(1.0, -k);
            for (long j = 0;; j++) {
                TSNode nd = node((2 * j + 1) * h);
                if (!(nd.xc >= min_c)) break;
                rows[k].push_back(nd);
            }
        }
    }
    // Returns the integral; *error and *L1 as Boost reports them. Throws never;
    // *finite is false if a sum became non-finite (Boost raises an evaluation error).
    template <class F>
    double integrate(F f, double a, double b, double tol, double* error, double* L1, bool* finite,
                     double abs_floor = 0.0, double value_cap = 0.0) const {
        double c = 0.5 * (a + b), d = 0.5 * (b - a);
        auto level_sum = [&](const std::vector<TSNode>& row, size_t first, double& s, double& as) {
            s = 0.0; as = 0.0;
            for (size_t i = first; i < row.size(); i++) {
                double off = d * row[i].xc;
                double xl = a + off, xr = b - off;
                if (xr != b) { double y = f(xr); s += y * row[i].w; as += std::fabs(y) * row[i].w; }
                if (xl != a) { double y = f(xl); s += y * row[i].w; as += std::fabs(y) * row[i].w; }
            }
        };
        double h = 1.0;
        double y0 = f(c);
        double I0 = half_pi() * y0, L0 = half_pi() * std::fabs(y0);
        double s, as;
        level_sum(rows[0], 1, s, as);
        I0 = (I0 + s) * h;
        L0 = (L0 + as) * h;
        double I1 = I0, L11 = L0, err = std::numeric_limits<double>::max();
        *finite = true;
        int lev = 0;
        for (int k = 1; k < max_levels; k++) {
            lev = k;
            I0 = I1; L0 = L11;
            I1 = 0.5 * I0; L11 = 0.5 * L0;
            h *= 0.5;
            level_sum(rows[k], 0, s, as);
            I1 += s * h;
            L11 += as * h;
            err = std::fabs(I0 - I1);
            if (!std::isfinite(I1)) { *finite = false; break; }
            if (err <= tol * L11) break;
            // successive estimates agree to the integrand's rounding noise, and the value is too small
            // for its error to be judged (see the success rule of integrate_error)
            if (err * d <= abs_floor && std::fabs(I1) * d <= value_cap) break;
        }
        *error = err * d;
        *L1 = L11 * d;
        static const bool trace = std::getenv("ORACLE_TRACE_QUAD") != nullptr;   // per-integral log
        if (trace)
            std::fprintf(stderr, "quad a=%.6e b=%.6e levels=%d I=%.6e err=%.3e\n", a, b, lev, I1 * d, err * d);
        return I1 * d;
    }
    static double half_pi() { return 1.5707963267948966; }
};

struct GaussLegendre {
    int n;
    std::vector<double> x, w;   // x[0] = 0 for odd n, then positive nodes
    explicit GaussLegendre(int nn) : n(nn) {
        std::vector<double> xs, ws;
        for (int i = 1; i <= n; i++) {
            double z = std::cos(M_PI * (i - 0.25) / (n + 0.5));
            double dp = 0.0;
            for (int it = 0; it < 100; it++) {
                double p0 = 1.0, p1 = z;
                for (int k = 2; k <= n; k++) {
                    double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1; p1 = p2;
                }
                dp = n * (z * p1 - p0) / (z * z - 1.0);
                double dz = p1 / dp;
                z -= dz;
                if (std::fabs(dz) < 1e-17) break;
            }
            {
                double p0 = 1.0, p1 = z;
                for (int k = 2; k <= n; k++) {
                    double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1; p1 = p2;
                }
                dp = n * (z * p1 - p0) / (z * z - 1.0);
            }
            if (z >= -1e-300) { xs.push_back(std::fabs(z) < 1e-15 ? 0.0 : z); ws.push_back(2.0 / ((1.0 - z * z) * dp * dp)); }
        }
        // order: centre first (odd n), then increasing positive nodes
        std::vector<int> idx(xs.size());
        for (size_t i = 0; i < idx.size(); i++) idx[i] = (int)i;
        std::sort(idx.begin(), idx.end(), [&](int a, int b) { return xs[a] < xs[b]; });
        for (int i : idx) { x.push_back(xs[i]); w.push_back(ws[i]); }
    }
    template <class F>
    double integrate(F f, double a, double b) const {
        double c = 0.5 * (a + b), d = 0.5 * (b - a);
        double s = 0.0;
        size_t first = 0;
        if (n & 1) { s = f(c) * w[0]; first = 1; }
        for (size_t i = first; i < x.size(); i++) s += (f(c + d * x[i]) + f(c - d * x[i])) * w[i];
        return s * d;
    }
};

const TanhSinh& tanh_sinh() { static TanhSinh ts; return ts; }
const GaussLegendre& gauss15() { static GaussLegendre g(15); return g; }
const GaussLegendre& gauss7() { static GaussLegendre g(7); return g; }

// integrateError (krylovTimeEvolver.cpp:584-618). Integrand h * |[exp(xH)e1]_{m-1}| (:591).
static double integrate_error(const KrylovSpace& ks, double a, double b, double h, int method,
                              double tolRate, double tol, double termination, bool& successful) {
    auto f = [&](double x) { return h * ks.last_abs(x); };
    double ret, error = 0.0, L1 = 0.0;
    bool success = true;
    if (method == 0) ret = gauss15().integrate(f, a, b);
    else if (method == 1) ret = gauss7().integrate(f, a, b);
    else {
        bool finite = true;
        // The integrand h |sum_k p_k e^{-i x lambda_k}| is bounded by h (sum_k |p_k| <= 1), so two
        // estimates that agree to eps h (b-a) agree to the rounding unit of that bound: refinement
        // stops there, but only below termination (b-a) tolRate, where the reference counts the
        // integral as successful whatever its error (:607-608).
        const double floor_ = DBL_EPSILON * h * (b - a);
        const double cap = termination * (b - a) * tolRate;
        ret = tanh_sinh().integrate(f, a, b, termination, &error, &L1, &finite, floor_, cap);
        if (!finite) { ret = tol * 10; error = 0.0; }        // Boost throws; caught at :602-605
        if (error / ret > termination && ret > termination * (b - a) * tolRate) success = false;  // :607-608
    }
    successful = successful && success;
    return ret;
}

// ----------------------------------------------------------------------------------
// The evolver: krylovTimeEvolver.cpp:26-84 (constructor), :261-512 (timeEvolve),
// :523-571 (arnoldiAlgorithm), :153-254 (findMaximalStepSize), :662-676 (return).
// ----------------------------------------------------------------------------------
struct Observable {
    int kind = 0;      // 0 sparse matrix, 1 vector
    Csr A;
    std::vector<cd> vec;
    std::vector<double> values;
};

struct Params {
    double t, samplingStep, expFactor, tol;
    int m, fastIntegration;
    long max_steps;        // 0 = unlimited (CPU-baseline timing hook)
    long stop_at_sample;   // -1 = never; emulates an observable throwing requestStopException
    long stop_obs;         // ... placed before observable stop_obs of the list (the ones before it
                           // record the stop sample, sample() evaluates them in order)
};

struct Result {
    double err = 0, evolvedTime = 0;
    long numSamples = 0, n_steps = 0, dim = 0, krylovDim = 0;
    int statusCode = 0;
};

struct Evolver {
    Params P;
    const Csr* H;
    long Hsize;
    int m;
    double matrixNorm;
    cd expFactor;
    long n_samples;
    std::vector<cd> currentVec, sampledState, tmp;
    std::vector<Observable>* obs;
    long index_samples = 0;
    double t_now = 0, t_sampling = 0, err = 0;
    long n_steps = 0;
    int statusCode = 0;
    bool stop_requested = false;
    std::vector<double> step_trace;

    // krylovSpMatrixObservable::expectation (krylovObservables.cpp:173-194) stores Re<v,Ov>;
    // krylovVectorObservable::expectation (:216-236) stores |<v,o>|^2.
    void sample() {
        const bool stop_here = P.stop_at_sample >= 0 && index_samples == P.stop_at_sample;
        for (size_t i = 0; i <= obs->size(); i++) {
            if (stop_here && (long)i == std::min<long>(P.stop_obs, (long)obs->size())) {
                stop_requested = true;
                return;
            }
            if (i == obs->size()) break;
            Observable& o = (*obs)[i];
            if (o.kind == 0) {
                spmv(o.A, cd(1.0, 0.0), sampledState.data(), tmp.data());
                cd v = zdotc(Hsize, sampledState.data(), tmp.data());
                o.values[index_samples] = v.real();
            } else {
                cd v = zdotc(Hsize, sampledState.data(), o.vec.data());
                o.values[index_samples] = std::norm(v);
            }
        }
        index_samples++;
    }

    // arnoldiAlgorithm (:523-571): Lanczos three-term recurrence for expFactor*H with
    // normalised columns of V; a[j] = H[j][j], b[j] = H[j+1][j] = -H[j][j+1].
    bool arnoldi(double tolRate, std::vector<cd>& V, std::vector<cd>& a, std::vector<double>& b,
                 double* hRet, int* mRet) {
        std::copy(currentVec.begin(), currentVec.end(), V.begin());
        for (int j = 0; j < m; j++) {
            cd* vj = &V[(size_t)j * Hsize];
            spmv(*H, expFactor, vj, tmp.data());
            if (j != 0) zaxpy(Hsize, cd(b[j - 1], 0.0), &V[(size_t)(j - 1) * Hsize], tmp.data());
            a[j] = zdotc(Hsize, vj, tmp.data());
            zaxpy(Hsize, -a[j], vj, tmp.data());
            double normy = dznrm2(Hsize, tmp.data());
            if (normy < tolRate) {          // lucky breakdown (:551-555)
                *mRet = j + 1;
                *hRet = normy;
                return true;
            }
            if (j + 1 != m) {
                b[j] = normy;
                double inv = 1.0 / normy;
                cd* vn = &V[(size_t)(j + 1) * Hsize];
                if (mkl::g_mkl) {                   // :563-564: zscal, then zcopy
                    const cd s(inv, 0.0);
                    mkl::g_mkl->zscal(Hsize, &s, tmp.data(), 1);
                    mkl::g_mkl->zcopy(Hsize, tmp.data(), 1, vn, 1);
                } else {
#pragma omp parallel for schedule(static)
                    for (long i = 0; i < Hsize; i++) vn[i] = tmp[i] * inv;
                }
            } else {
                *hRet = normy;
            }
        }
        *mRet = m;
        return false;
    }

    // findMaximalStepSize (:153-254)
    int find_step(const KrylovSpace& ks, double h, double tolRate, double t_stepSuggestion,
                  double t_step_max, int n_substeps, double numericalErrorEstimate, bool increaseStep,
                  int methodLong, int methodShort, double* t_stepRet, cd* wRet, double* err_stepRet) {
        const unsigned GO_MAX = 20;
        unsigned nbReductions = 0;
        double termination = 1e-3;
        double t_step = t_stepSuggestion;
        int code = 0;
        bool okInt = true;
        bool skipSubsteps = false;
        if (t_step > t_step_max) { t_step = t_step_max; skipSubsteps = true; }
        double err_step = integrate_error(ks, 0, t_step, h, methodLong, tolRate, P.tol, termination, okInt);
        if (increaseStep && !skipSubsteps) {
            double err_new = err_step, t_new = t_step;
            while (err_new < tolRate * t_new && t_step < t_step_max) {
                t_step = t_new;
                err_step = err_new;
                if (t_step >= t_step_max) { skipSubsteps = true; break; }
                t_new *= 2;
                if (t_new > t_step_max) t_new = t_step_max;
                err_new = integrate_error(ks, 0, t_new, h, methodLong, tolRate, P.tol, termination, okInt);
            }
        }
        while (err_step > tolRate * t_step) {
            nbReductions++;
            t_step = t_step / 2.0;
            err_step = integrate_error(ks, 0, t_step, h, methodLong, tolRate, P.tol, termination, okInt);
            if (nbReductions == GO_MAX) return -1;   // reference: FATAL + exit(1) (:203-207)
            skipSubsteps = false;
        }
        double s = t_step / n_substeps;
        int n_s = 0;
        double deltaError = 0;
        if (!skipSubsteps) {
            while (err_step + deltaError < tolRate * (t_step + n_s * s)) {
                err_step += deltaError;
                n_s++;
                if (t_step + (n_s - 1) * s >= t_step_max) break;
                deltaError = integrate_error(ks, t_step + (n_s - 1) * s, t_step + n_s * s, h, methodShort,
                                             tolRate, P.tol, termination, okInt);
            }
            t_step += (n_s - 1) * s;
        }
        if (err_step < numericalErrorEstimate) {
            if (!skipSubsteps || err_step > 0.1 * tolRate * t_step) code += 1;
        }
        if (!okInt) code += 20;
        *t_stepRet = t_step;
        *err_stepRet = err_step;
        ks.exp_e1(t_step, wRet);
        return code;
    }

    // timeEvolve (:261-512). Returns 0, or -1 where the reference exits with FATAL.
    int run(Result& R) {
        const int N_SUBSTEPS = 50;
        const double INITIAL_STEP_FRACTION = 0.97;
        int methodLong, methodShort;
        if (P.fastIntegration) { methodLong = 0; methodShort = 1; }
        else { methodLong = methodShort = 2; }
        double t_step = 1.0 / 10.0;
        double tolRate = P.tol / P.t;
        double numericalErrorEstimate = Hsize * matrixNorm * std::numeric_limits<double>::epsilon();
        double numericalErrorEstimateTotal = 0;
        bool hbd = false;
        int m_hbd = 0;
        int errorCodeFindSubstep = 0;
        int nbErrRoundOff = 0, nbErrInt = 0, nbErrors = 0;

        sample();
        if (stop_requested) { finish(R); return 0; }

        std::vector<cd> V((size_t)Hsize * m);
        std::vector<cd> a(m), w(m), y(m);
        std::vector<double> b(m, 0.0);
        double h = 0;
        KrylovSpace ks;

        while (index_samples < n_samples) {
            if (P.max_steps > 0 && n_steps >= P.max_steps) break;
            n_steps++;
            double err_step = 0;
            hbd = arnoldi(tolRate, V, a, b, &h, &m_hbd);
            if (hbd) {                                  // :342-363
                t_step = P.t - t_now;
                m = m_hbd;
                statusCode = 1;
            }
            if (ks.factor(m, a.data(), b.data()) != 0) return -1;   // :365-370
            if (!hbd) {
                double s0 = INITIAL_STEP_FRACTION * t_step;
                errorCodeFindSubstep = find_step(ks, h, tolRate, s0, P.t - t_now, N_SUBSTEPS,
                                                 numericalErrorEstimate, t_now == 0, methodLong, methodShort,
                                                 &t_step, w.data(), &err_step);
                if (errorCodeFindSubstep < 0) return -1;
                gemv_tall(Hsize, m, V.data(), w.data(), currentVec.data());   // :382
            }
            switch (errorCodeFindSubstep) {            // :386-399
                case 0: break;
                case 1: nbErrRoundOff++; break;
                case 20: nbErrInt++; break;
                case 21: nbErrInt++; nbErrRoundOff++; break;
            }
            step_trace.push_back(t_step);
            // sampling (:405-437)
            while (index_samples < n_samples &&
                   (t_sampling + P.samplingStep <= t_now + t_step ||
                    (index_samples == n_samples - 1 &&
                     std::abs(t_now + t_step - P.t) <= std::numeric_limits<double>::epsilon()))) {
                t_sampling += P.samplingStep;
                ks.exp_e1(t_sampling - t_now, y.data());
                gemv_tall(Hsize, m, V.data(), y.data(), sampledState.data());
                sample();
                if (stop_requested) {
                    statusCode = 3;
                    finish(R);
                    return 0;
                }
            }
            t_now += t_step;
            err += err_step;
        }
        // :446-452
        if (std::abs(dznrm2(Hsize, sampledState.data()) - 1.0) > P.tol) {
            nbErrors++;
            statusCode = 30;
        }
        // :459-504
        numericalErrorEstimateTotal = numericalErrorEstimate * n_steps;
        if (nbErrRoundOff != 0) { nbErrors++; statusCode = 10; }
        if (nbErrInt != 0) { nbErrors++; statusCode = 20; }
        if (numericalErrorEstimateTotal > err && nbErrRoundOff == 0) {
            if (numericalErrorEstimateTotal > P.tol) { statusCode = 11; nbErrors++; }
            else if (statusCode != 1) statusCode = 2;
        }
        if (nbErrors > 1) statusCode = 100;
        finish(R);
        return 0;
    }

    void finish(Result& R) {                        // generateReturn (:662-676)
        R.err = err;
        R.n_steps = n_steps;
        R.dim = Hsize;
        R.krylovDim = m;
        R.evolvedTime = t_now;
        R.numSamples = index_samples;
        R.statusCode = statusCode;
    }
};

}  // namespace

// ==================================================================================
// C ABI for ctypes (tests/, bench.py cpu_baseline). All indices are int64.
// ==================================================================================
extern "C" {

void or_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int or_get_threads() { return omp_get_max_threads(); }

void* or_basis_new(int N, int K, int nbQudits, int capacity, int quant) {
    Basis* b = make_basis(N, K, nbQudits, capacity, quant);
    return b;
}
void* or_tensor_basis_new(int N1, int K1, int N2, int K2, int C2) { return make_tensor_basis(N1, K1, N2, K2, C2); }
long or_basis_size(void* h) { return ((Basis*)h)->D; }
int or_basis_modes(void* h) { return ((Basis*)h)->K; }
void or_basis_states(void* h, int* out) {
    Basis* b = (Basis*)h;
    std::memcpy(out, b->occ.data(), sizeof(int) * b->occ.size());
}
void or_basis_index(void* h) { ((Basis*)h)->build_table(); }
long or_basis_find(void* h, const int* e) {
    Basis* b = (Basis*)h;
    if (!b->indexed) b->build_table();
    return b->find(e);
}
void or_basis_free(void* h) { delete (Basis*)h; }

static std::vector<Term> unpack_terms(int n_terms, const double* coef, const int* n_ops, const int* ops,
                                      const int* modes, const int* caps) {
    std::vector<Term> terms(n_terms);
    long o = 0;
    for (int t = 0; t < n_terms; t++) {
        terms[t].coef = cd(coef[2 * t], coef[2 * t + 1]);
        for (int i = 0; i < n_ops[t]; i++, o++) {
            terms[t].ops.push_back(ops[o]);
            terms[t].modes.push_back(modes[o]);
            terms[t].caps.push_back(caps[o]);
        }
    }
    return terms;
}

void* or_build_operator(void* basis, int n_terms, const double* coef, const int* n_ops, const int* ops,
                        const int* modes, const int* caps) {
    Basis* b = (Basis*)basis;
    if (!b->indexed) b->build_table();
    return build_operator(b, unpack_terms(n_terms, coef, n_ops, ops, modes, caps));
}
long or_coo_nnz(void* h) { return (long)((Coo*)h)->val.size(); }
void or_coo_copy(void* h, double* vals, long* cols, long* rows) {
    Coo* m = (Coo*)h;
    for (size_t i = 0; i < m->val.size(); i++) {
        vals[2 * i] = m->val[i].real();
        vals[2 * i + 1] = m->val[i].imag();
        cols[i] = m->col[i];
        rows[i] = m->row[i];
    }
}
void or_coo_free(void* h) { delete (Coo*)h; }

// Rows of createMatrix for selected basis indices, with image states found through a
// caller-supplied flat basis and rank function is not needed: uses the hash table.
// Writes up to cap entries per row; returns total written into counts.
void or_build_rows(void* basis, int n_terms, const double* coef, const int* n_ops, const int* ops,
                   const int* modes, const int* caps, long n_rows, const long* row_ids, int cap,
                   int* counts, long* cols, double* vals) {
    Basis* b = (Basis*)basis;
    if (!b->indexed) b->build_table();
    std::vector<Term> terms = unpack_terms(n_terms, coef, n_ops, ops, modes, caps);
    auto finder = [b](const int* e) { return b->find(e); };
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < n_rows; i++) {
        std::vector<std::pair<long, cd>> out;
        build_row(terms, &b->occ[(size_t)row_ids[i] * b->K], b->K, finder, out);
        int c = (int)std::min<size_t>(out.size(), (size_t)cap);
        counts[i] = (int)out.size();
        for (int k = 0; k < c; k++) {
            cols[i * cap + k] = out[k].first;
            vals[2 * (i * cap + k)] = out[k].second.real();
            vals[2 * (i * cap + k) + 1] = out[k].second.imag();
        }
    }
}

double or_norm1(long ncols, long nnz, const double* v, const long* col) { return norm1_coo(ncols, nnz, v, col); }

void or_spmv(long n, long nnz, const double* v, const long* col, const long* row, double alpha_re,
             double alpha_im, const double* x, double* y) {
    Csr A = coo_to_csr(n, n, nnz, v, col, row);
    spmv(A, cd(alpha_re, alpha_im), (const cd*)x, (cd*)y);
}

// Tridiagonal eigen-solver and Krylov exponential, exposed for unit tests.
int or_tridiag_eig(int n, const double* a_im, const double* b, double* lam, double* Q) {
    std::vector<double> d(n), e(n, 0.0);
    for (int j = 0; j < n; j++) d[j] = -a_im[j];
    for (int j = 0; j + 1 < n; j++) e[j] = b[j];
    int info = tridiag_ql(n, d.data(), e.data(), Q);
    for (int j = 0; j < n; j++) lam[j] = d[j];
    return info;
}
// a: m complex (interleaved) Lanczos diagonal, b: m-1 off-diagonals; out: m complex
int or_exp_krylov(int m, const double* a, const double* b, double x, double* out) {
    KrylovSpace ks;
    int info = ks.factor(m, (const cd*)a, b);
    ks.exp_e1(x, (cd*)out);
    return info;
}
// Error integral on [lo, hi] for the Krylov space (a, b) with last norm h.
// out: [value, error, L1] for tanh-sinh; value only for Gauss. Returns success flag.
int or_integrate_error(int m, const double* a, const double* b, double h, double lo, double hi, int method,
                       double tolRate, double tol, double* out) {
    KrylovSpace ks;
    ks.factor(m, (const cd*)a, b);
    bool ok = true;
    double v = integrate_error(ks, lo, hi, h, method, tolRate, tol, 1e-3, ok);
    out[0] = v;
    if (method == 2) {
        auto f = [&](double x) { return h * ks.last_abs(x); };
        double e, L1;
        bool fin;
        tanh_sinh().integrate(f, lo, hi, 1e-3, &e, &L1, &fin);
        out[1] = e;
        out[2] = L1;
    }
    return ok ? 1 : 0;
}
// Generic tanh-sinh on f(x) = sum_k c_k x^k (for pinning the quadrature).
double or_tanh_sinh_poly(int deg, const double* c, double a, double b, double tol, double* err, double* L1) {
    auto f = [&](double x) { double s = 0; for (int k = deg; k >= 0; k--) s = s * x + c[k]; return s; };
    bool fin;
    return tanh_sinh().integrate(f, a, b, tol, err, L1, &fin);
}
double or_gauss_poly(int n, int deg, const double* c, double a, double b) {
    auto f = [&](double x) { double s = 0; for (int k = deg; k >= 0; k--) s = s * x + c[k]; return s; };
    return (n == 15 ? gauss15() : gauss7()).integrate(f, a, b);
}
int or_tanh_sinh_rows(int level) { return (int)tanh_sinh().rows[level].size(); }

// Diagonal observables given as dense vectors of their diagonal (CPU-baseline entry point).
static Csr diag_csr(long n, const double* d) {
    Csr A;
    A.n = A.ncols = n;
    A.ptr.resize(n + 1);
    A.col.resize(n);
    A.val.resize(n);
    for (long r = 0; r <= n; r++) A.ptr[r] = r;
    for (long r = 0; r < n; r++) { A.col[r] = r; A.val[r] = cd(d[r], 0.0); }
    return A;
}

// Full time evolution. H and sparse observables given as COO (int64 indices). stop_at_sample
// >= 0 emulates a host observable throwing requestStopException at that sample, placed before
// observable stop_obs of the list.
// obs_kind[i]: 0 sparse (obs_nnz[i] entries taken in turn from obs_vals/cols/rows),
// 1 vector (D complex entries taken in turn from obs_vecs).
// obs_out: n_obs x n_samples. result: [err, evolvedTime, numSamples, n_steps, dim,
// krylovDim, statusCode, n_samples_alloc]. trace (optional, may be NULL): t_step per
// Krylov step, up to trace_cap entries. Returns 0, or -1 where the reference would exit(1).
int or_evolve(long n, long nnz, const double* hv, const long* hc, const long* hr, const double* psi0,
              double t, double samplingStep, double expFactor, double tol, int m, int fastIntegration,
              long max_steps, long stop_at_sample, long stop_obs, int n_obs, const int* obs_kind, const long* obs_nnz,
              const double* obs_vals, const long* obs_cols, const long* obs_rows, const double* obs_vecs,
              double* obs_out, double* state_out, double* result, double* trace, long trace_cap) {
    Csr H = coo_to_csr(n, n, nnz, hv, hc, hr);
    Evolver ev;
    ev.P = Params{t, samplingStep, expFactor, tol, m, fastIntegration, max_steps, stop_at_sample, stop_obs};
    ev.H = &H;
    ev.Hsize = n;
    ev.matrixNorm = norm1_coo(n, nnz, hv, hc);
    ev.m = (int)std::min<long>(m, n);
    ev.expFactor = cd(0.0, -1.0 * expFactor);
    if (n == 0) return -1;
    if (std::abs(dznrm2(n, (const cd*)psi0) - 1.0) > tol) return -1;   // :45-48
    ev.n_samples = (long)std::floor(t / samplingStep) + 1;             // :57
    ev.currentVec.assign((const cd*)psi0, (const cd*)psi0 + n);
    ev.sampledState = ev.currentVec;
    ev.tmp.assign(n, cd(0.0, 0.0));
    std::vector<Observable> obs(n_obs);
    long ov = 0, ovv = 0;
    for (int i = 0; i < n_obs; i++) {
        obs[i].kind = obs_kind[i];
        if (obs_kind[i] == 0) {
            obs[i].A = coo_to_csr(n, n, obs_nnz[i], obs_vals + 2 * ov, obs_cols + ov, obs_rows + ov);
            ov += obs_nnz[i];
        } else {
            obs[i].vec.assign((const cd*)obs_vecs + ovv, (const cd*)obs_vecs + ovv + n);
            ovv += n;
        }
        obs[i].values.assign(ev.n_samples, 0.0);
    }
    ev.obs = &obs;
    if (mkl_attach(H) != 0) return -2;
    for (auto& o : obs)
        if (o.kind == 0 && mkl_attach(o.A) != 0) return -2;
    Result R;
    int rc = ev.run(R);
    if (rc != 0) return rc;
    for (int i = 0; i < n_obs; i++)
        for (long s = 0; s < ev.n_samples; s++) obs_out[i * ev.n_samples + s] = obs[i].values[s];
    std::memcpy(state_out, ev.sampledState.data(), sizeof(cd) * n);
    result[0] = R.err; result[1] = R.evolvedTime; result[2] = (double)R.numSamples;
    result[3] = (double)R.n_steps; result[4] = (double)R.dim; result[5] = (double)R.krylovDim;
    result[6] = (double)R.statusCode; result[7] = (double)ev.n_samples;
    if (trace)
        for (size_t i = 0; i < ev.step_trace.size() && (long)i < trace_cap; i++) trace[i] = ev.step_trace[i];
    return 0;
}

// Time evolution with H given as CSR (int64 row pointers, int32 columns) and diagonal
// observables as dense diagonals (n_obs x n): the CPU baseline on a large workload, where a
// COO sort would dominate. max_steps bounds the number of Krylov steps (0 = all).
// result: [err, evolvedTime, numSamples, n_steps, dim, krylovDim, statusCode, n_samples,
// seconds of timeEvolve (the run loop: samples, Lanczos, step control, state updates),
// seconds of the set-up before it (copies, norm1, and with MKL the handles' initialize)].
int or_evolve_csr(long n, const long* ptr, const int* col, const double* vals, const double* psi0, double t,
                  double samplingStep, double expFactor, double tol, int m, int fastIntegration, long max_steps,
                  int n_obs, const double* diags, double* obs_out, double* state_out, double* result) {
    const double t_setup = omp_get_wtime();
    Csr H;
    H.n = H.ncols = n;
    long nnz = ptr[n];
    H.ptr.assign(ptr, ptr + n + 1);
    H.col.resize(nnz);
    H.val.resize(nnz);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < nnz; k++) {
        H.col[k] = col[k];
        H.val[k] = cd(vals[2 * k], vals[2 * k + 1]);
    }
    std::vector<double> colVal(n, 0.0);
    for (long k = 0; k < nnz; k++) colVal[H.col[k]] += std::abs(H.val[k]);
    Evolver ev;
    ev.P = Params{t, samplingStep, expFactor, tol, m, fastIntegration, max_steps, -1, 0};
    ev.H = &H;
    ev.Hsize = n;
    ev.matrixNorm = *std::max_element(colVal.begin(), colVal.end());
    ev.m = (int)std::min<long>(m, n);
    ev.expFactor = cd(0.0, -1.0 * expFactor);
    if (std::abs(dznrm2(n, (const cd*)psi0) - 1.0) > tol) return -1;
    ev.n_samples = (long)std::floor(t / samplingStep) + 1;
    ev.currentVec.assign((const cd*)psi0, (const cd*)psi0 + n);
    ev.sampledState = ev.currentVec;
    ev.tmp.assign(n, cd(0.0, 0.0));
    std::vector<Observable> obs(n_obs);
    for (int i = 0; i < n_obs; i++) {
        obs[i].kind = 0;
        obs[i].A = diag_csr(n, diags + (size_t)i * n);
        obs[i].values.assign(ev.n_samples, 0.0);
    }
    ev.obs = &obs;
    if (mkl_attach(H) != 0) return -2;
    for (auto& o : obs)
        if (mkl_attach(o.A) != 0) return -2;
    Result R;
    const double t_run = omp_get_wtime();
    int rc = ev.run(R);
    const double t_end = omp_get_wtime();
    if (rc != 0) return rc;
    for (int i = 0; i < n_obs; i++)
        for (long s = 0; s < ev.n_samples; s++) obs_out[i * ev.n_samples + s] = obs[i].values[s];
    if (state_out) std::memcpy(state_out, ev.sampledState.data(), sizeof(cd) * n);
    result[0] = R.err; result[1] = R.evolvedTime; result[2] = (double)R.numSamples;
    result[3] = (double)R.n_steps; result[4] = (double)R.dim; result[5] = (double)R.krylovDim;
    result[6] = (double)R.statusCode; result[7] = (double)ev.n_samples;
    result[8] = t_end - t_run; result[9] = t_run - t_setup;
    return 0;
}

// MKL backend switch (CPU baseline): 0 when libmkl_rt was opened and every entry resolved
// (ILP64 interface, GNU OpenMP threading, `threads` threads; 0 = OpenMP's count), else the
// loop kernels stay in use and the code says which step failed (1 no library, 2 missing
// entry, 3 ILP64 refused). or_mkl_disable returns to the loop kernels (the library stays
// loaded; a later or_mkl_enable switches it back on).
int or_mkl_enable(const char* path, int threads) { return mkl::enable(path, threads); }
void or_mkl_disable() { mkl::g_mkl = nullptr; }
// The MKL version string ("" while the backend is off).
void or_mkl_version(char* buf, int len) {
    if (len <= 0) return;
    std::snprintf(buf, (size_t)len, "%s", mkl::g_mkl ? mkl::g_api.ver : "");
}

}  // extern "C"
