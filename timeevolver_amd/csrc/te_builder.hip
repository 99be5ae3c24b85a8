// Bit-exact device builder of the occupation-number basis and of operator matrices.
//
// Replaces basis::basis (helper/Basis.cpp:37-120), tensorBasis (:131-182), the hash-table
// lookup (helper/Basis.h:284, used at helper/hamiltonian.cpp:360-369) and
// Hamiltonian::createMatrix (helper/hamiltonian.cpp:318-401, with creationOperator :410-434,
// annihilationOperator :441-456 and scalarMultiplication :308-311).
//
// Basis: the reference grows a list level by level: starting from [0-vector], for
// k = K..2 every state of the list spawns, in list order, copies with e[k-1] = quant,
// 2 quant, ..., upper that are appended behind the whole list. On the device each level is
// a child count, an exclusive scan and a scatter; the scans are kept because they rank any
// occupation vector without a hash table:
//   r_K = 0;  r_p = r_{p+1} if e_p = 0, else |A_{p+1}| + off_{p+1}[r_{p+1}] + e_p/quant - 1,
// valid iff 1 <= e_p/quant <= (children of r_{p+1}); finally e_0 = N - sum and the
// all-qudit filter is a prefix over the kept flags.
//
// Operator rows: one thread per basis state applies the terms in list order with the
// reference's exact multiplication sequence (compiled with -ffp-contract=off; square roots
// come from a host table of correctly rounded sqrt(k)), accumulates equal images in term
// order (std::map insert-then-+=), sorts by column and drops |v| < DBL_EPSILON. A count
// pass sizes the rows, a fill pass writes them.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <climits>
#include <cstdint>
#include <cstring>
#include <vector>

#include "te_internal.h"

namespace te {

constexpr int kMaxModes = 64;

struct FactorDev {
    int N, K, nbQ, cap, quant;
    int64_t size[kMaxModes + 1];       // |A_p|, p = 1..K (|A_K| = 1)
    const int64_t* off[kMaxModes + 1]; // off[p], p = 2..K: exclusive scan of child counts over A_p
    const int64_t* keep;               // prefix over A_1 when the all-qudit filter applies
    int64_t D;
};

struct BasisDev {
    int K, K1, nf;
    int64_t D;
    const int* occ;                    // D x K
    FactorDev f[2];
};

}  // namespace te

struct te_basis {
    te_ctx* ctx = nullptr;
    int K = 0;
    int64_t D = 0;
    int max_occ = 0;                   // largest occupation of any mode in any state
    int* occ = nullptr;
    te::BasisDev dev{};
    std::vector<void*> allocations;
    // Lexicographic order of the states (computed on first use by an operator build): the row
    // order of the operators' value-dictionary copies, new -> old. Null if not applicable.
    mutable int* lex_perm = nullptr;
    mutable bool lex_done = false;
};

namespace te {

// ---------------------------------------------------------------------------------------
// ranking
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ int64_t rank_factor(const FactorDev& f, const int* e) {
    const int K = f.K;
    int tot = 0;
    for (int p = 1; p < K; p++) tot += e[p];
    if (e[0] != f.N - tot) return -1;                  // particle number (e_0 = N - sum)
    int64_t r = 0;
    for (int p = K - 1; p >= 1; p--) {
        const int x = e[p];
        if (x == 0) continue;
        if (x < 0 || x % f.quant != 0) return -1;
        const int64_t c = x / f.quant;
        const int64_t o0 = f.off[p + 1][r], o1 = f.off[p + 1][r + 1];
        if (c > o1 - o0) return -1;
        r = f.size[p + 1] + o0 + c - 1;
    }
    if (f.keep) {
        const int64_t k0 = f.keep[r], k1 = f.keep[r + 1];
        if (k1 == k0) return -1;
        r = k0;
    }
    return r;
}

__device__ __forceinline__ int64_t rank_basis(const BasisDev& B, const int* e) {
    if (B.nf == 1) return rank_factor(B.f[0], e);
    const int64_t r1 = rank_factor(B.f[0], e);
    if (r1 < 0) return -1;
    const int64_t r2 = rank_factor(B.f[1], e + B.K1);
    if (r2 < 0) return -1;
    return r1 * B.f[1].D + r2;
}

__global__ void find_kernel(BasisDev B, const int* occ, int64_t count, int64_t* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int e[kMaxModes];
    for (int k = 0; k < B.K; k++) e[k] = occ[i * B.K + k];
    out[i] = rank_basis(B, e);
}

// ---------------------------------------------------------------------------------------
// basis levels
// ---------------------------------------------------------------------------------------
__global__ void child_count_kernel(const int* sums, int64_t n, int N, int cap, int quant, int qudit, int64_t* cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        cnt[i] = 0;
        return;
    }
    int upper = N - sums[i];
    if (qudit) upper = upper < cap ? upper : cap;
    cnt[i] = upper >= quant ? upper / quant : 0;
}

__global__ void expand_kernel(const int* sums, int64_t n, const int64_t* off, int quant, int* par, int* occ,
                              int* sums_new) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = sums[i];
    sums_new[i] = s;
    const int64_t o0 = off[i], o1 = off[i + 1];
    for (int64_t j = o0; j < o1; j++) {
        const int o = (int)(j - o0 + 1) * quant;
        par[j] = (int)i;
        occ[j] = o;
        sums_new[n + j] = s + o;
    }
}

struct LevelPtrs {
    const int* par[kMaxModes + 1];
    const int* occ[kMaxModes + 1];
    int64_t size[kMaxModes + 1];
};

// Final list A_1 -> full occupation vectors; keep flags for the all-qudit filter.
__global__ void materialise_kernel(LevelPtrs L, int K, int N, int cap, int filter, int64_t n, int* out,
                                   int64_t* keep) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        if (keep) keep[n] = 0;
        return;
    }
    int64_t idx = i;
    int tot = 0;
    for (int p = 1; p < K; p++) {
        int x = 0;
        if (idx >= L.size[p + 1]) {
            const int64_t j = idx - L.size[p + 1];
            x = L.occ[p][j];
            idx = L.par[p][j];
        }
        out[i * K + p] = x;
        tot += x;
    }
    const int missing = N - tot;
    out[i * K] = missing;
    if (keep) keep[i] = (!filter || missing <= cap) ? 1 : 0;
}

__global__ void compact_kernel(const int* in, const int64_t* keep_scan, int64_t n, int K, int* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (keep_scan[i + 1] == keep_scan[i]) return;
    const int64_t d = keep_scan[i];
    for (int k = 0; k < K; k++) out[d * K + k] = in[i * K + k];
}

__global__ void tensor_kernel(const int* a, int64_t na, int Ka, const int* b, int64_t nb, int Kb, int* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na * nb) return;
    const int64_t ia = i / nb, ib = i % nb;
    const int K = Ka + Kb;
    for (int k = 0; k < Ka; k++) out[i * K + k] = a[ia * Ka + k];
    for (int k = 0; k < Kb; k++) out[i * K + Ka + k] = b[ib * Kb + k];
}

static unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }

static int exclusive_scan(te_ctx* ctx, const int64_t* in, int64_t* out, int64_t n) {
    size_t bytes = 0;
    TE_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, ctx->stream));
    void* tmp = nullptr;
    TE_HIP(ctx, hipMalloc(&tmp, std::max<size_t>(bytes, 1)));
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, ctx->stream);
    hipStreamSynchronize(ctx->stream);
    hipFree(tmp);
    TE_HIP(ctx, e);
    return TE_OK;
}

// Builds one factor basis(N, K, nbQ, cap, quant); fills f and writes the states to *occ_out.
static int build_factor(te_ctx* ctx, te_basis* B, int N, int K, int nbQ, int cap, int quant, FactorDev& f,
                        int** occ_out) {
    hipStream_t s = ctx->stream;
    f = FactorDev{};
    f.N = N;
    f.K = K;
    f.nbQ = nbQ;
    f.cap = cap;
    f.quant = quant;
    int* sums = nullptr;
    TE_HIP(ctx, hipMalloc(&sums, sizeof(int)));
    TE_HIP(ctx, hipMemsetAsync(sums, 0, sizeof(int), s));
    int64_t size = 1;
    LevelPtrs L{};
    std::vector<void*> level_tmp;
    f.size[K] = 1;
    for (int p = K - 1; p >= 1; p--) {               // writes mode p (k = p + 1 in the reference)
        const int qudit = (p + 1 > K - nbQ) ? 1 : 0;
        int64_t *cnt = nullptr, *off = nullptr;
        TE_HIP(ctx, hipMalloc(&cnt, sizeof(int64_t) * (size + 1)));
        TE_HIP(ctx, hipMalloc(&off, sizeof(int64_t) * (size + 1)));
        hipLaunchKernelGGL(child_count_kernel, dim3(blocks_for(size + 1)), dim3(256), 0, s, sums, size, N, cap, quant,
                           qudit, cnt);
        int rc = exclusive_scan(ctx, cnt, off, size + 1);
        if (rc != TE_OK) return rc;
        int64_t total = 0;
        TE_HIP(ctx, hipMemcpyAsync(&total, off + size, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        TE_HIP(ctx, hipStreamSynchronize(s));
        if (size + total >= (int64_t)INT32_MAX) return fail(ctx, TE_ERR_UNSUPPORTED, "basis exceeds 2^31 states");
        int *par = nullptr, *occ = nullptr, *sums_new = nullptr;
        TE_HIP(ctx, hipMalloc(&par, sizeof(int) * std::max<int64_t>(total, 1)));
        TE_HIP(ctx, hipMalloc(&occ, sizeof(int) * std::max<int64_t>(total, 1)));
        TE_HIP(ctx, hipMalloc(&sums_new, sizeof(int) * (size + total)));
        hipLaunchKernelGGL(expand_kernel, dim3(blocks_for(size)), dim3(256), 0, s, sums, size, off, quant, par, occ,
                           sums_new);
        TE_HIP(ctx, hipGetLastError());
        hipFree(cnt);
        TE_HIP(ctx, hipStreamSynchronize(s));
        hipFree(sums);
        sums = sums_new;
        B->allocations.push_back(off);
        f.off[p + 1] = off;
        L.par[p] = par;
        L.occ[p] = occ;
        level_tmp.push_back(par);
        level_tmp.push_back(occ);
        size += total;
        f.size[p] = size;
    }
    if (K == 1) f.size[1] = 1;
    for (int p = 1; p <= K; p++) L.size[p] = f.size[p];
    const int64_t n1 = f.size[1];
    const int filter = (K == nbQ) ? 1 : 0;
    int* full = nullptr;
    TE_HIP(ctx, hipMalloc(&full, sizeof(int) * n1 * K));
    int64_t *keep = nullptr, *keep_scan = nullptr;
    if (filter) {
        TE_HIP(ctx, hipMalloc(&keep, sizeof(int64_t) * (n1 + 1)));
        TE_HIP(ctx, hipMalloc(&keep_scan, sizeof(int64_t) * (n1 + 1)));
    }
    hipLaunchKernelGGL(materialise_kernel, dim3(blocks_for(n1 + 1)), dim3(256), 0, s, L, K, N, cap, filter, n1, full,
                       keep);
    TE_HIP(ctx, hipGetLastError());
    if (filter) {
        int rc = exclusive_scan(ctx, keep, keep_scan, n1 + 1);
        if (rc != TE_OK) return rc;
        int64_t D = 0;
        TE_HIP(ctx, hipMemcpyAsync(&D, keep_scan + n1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        TE_HIP(ctx, hipStreamSynchronize(s));
        int* comp = nullptr;
        TE_HIP(ctx, hipMalloc(&comp, sizeof(int) * std::max<int64_t>(D, 1) * K));
        hipLaunchKernelGGL(compact_kernel, dim3(blocks_for(n1)), dim3(256), 0, s, full, keep_scan, n1, K, comp);
        TE_HIP(ctx, hipStreamSynchronize(s));
        hipFree(full);
        hipFree(keep);
        full = comp;
        f.keep = keep_scan;
        B->allocations.push_back(keep_scan);
        f.D = D;
    } else {
        f.D = n1;
    }
    TE_HIP(ctx, hipStreamSynchronize(s));
    hipFree(sums);
    for (void* p : level_tmp) hipFree(p);
    *occ_out = full;
    return TE_OK;
}

// ---------------------------------------------------------------------------------------
// operator rows
// ---------------------------------------------------------------------------------------
struct TermsDev {
    int n_terms;
    int max_entries;       // capacity of a row's entry list (= n_terms)
    const int* term_off;   // n_terms + 1 offsets into ops/modes/caps
    const int* ops;
    const int* modes;
    const int* caps;
    const double2* coef;
    const double* sqrt_tab;
    int sqrt_n;
};

// Per-thread workspace: vals[max_entries] (double2), cols[max_entries] (int64), e[K], w[K].
__device__ __forceinline__ size_t ws_bytes(int max_entries, int K) {
    return ((size_t)max_entries * (16 + 8) + (size_t)2 * K * 4 + 15) / 16 * 16;
}

// Builds row j's accumulated entries; returns the number of distinct columns.
__device__ int build_row(const BasisDev& B, const TermsDev& T, int64_t j, double2* vals, int64_t* cols, int* e,
                         int* w) {
    const int K = B.K;
    for (int k = 0; k < K; k++) e[k] = B.occ[j * K + k];
    int n_ent = 0;
    for (int t = 0; t < T.n_terms; t++) {
        for (int k = 0; k < K; k++) w[k] = e[k];
        double cr = 1.0, ci = 0.0;
        for (int q = T.term_off[t]; q < T.term_off[t + 1]; q++) {
            const int mode = T.modes[q];
            const int n = w[mode];
            if (T.ops[q] == 1) {                                   // creationOperator (:410-434)
                const int c = T.caps[q];
                if (c != 0 && n + 1 > c) {
                    cr = 0.0;
                    ci = 0.0;
                } else {
                    const double sq = T.sqrt_tab[n + 1];
                    cr = cr * sq;
                    ci = ci * sq;
                    w[mode] = n + 1;
                }
            } else if (T.ops[q] == -1) {                           // annihilationOperator (:441-456)
                if (n == 0) {
                    cr = 0.0;
                    ci = 0.0;
                } else {
                    const double sq = T.sqrt_tab[n];
                    cr = cr * sq;
                    ci = ci * sq;
                    w[mode] = n - 1;
                }
            }
        }
        const double2 tc = T.coef[t];                             // scalarMultiplication (:308-311)
        const double vr = cr * tc.x - ci * tc.y;
        const double vi = cr * tc.y + ci * tc.x;
        bool same = true;                                          // image == source: its own index
        for (int k = 0; k < K; k++) same = same && (w[k] == e[k]);
        const int64_t idx = same ? j : rank_basis(B, w);
        if (idx < 0) continue;                                     // not in the hash table (:360-369)
        int k = 0;
        while (k < n_ent && cols[k] != idx) k++;
        if (k < n_ent) {                                           // += on an existing map entry
            vals[k].x = vals[k].x + vr;
            vals[k].y = vals[k].y + vi;
        } else {                                                   // first insertion
            cols[n_ent] = idx;
            vals[n_ent] = make_double2(vr, vi);
            n_ent++;
        }
    }
    return n_ent;
}

template <bool FILL>
__global__ void operator_rows_kernel(BasisDev B, TermsDev T, char* ws_global, int64_t* counts, const int64_t* row_ptr,
                                     int* col_out, double2* val_out) {
    extern __shared__ __align__(16) char smem[];
    const size_t per = ws_bytes(T.max_entries, B.K);
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    char* base = ws_global ? ws_global + (size_t)gtid * per : smem + (size_t)threadIdx.x * per;
    double2* vals = (double2*)base;
    int64_t* cols = (int64_t*)(base + (size_t)T.max_entries * 16);
    int* e = (int*)(base + (size_t)T.max_entries * 24);
    int* w = e + B.K;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = gtid; j < B.D; j += stride) {
        const int n_ent = build_row(B, T, j, vals, cols, e, w);
        if (!FILL) {
            int c = 0;
            for (int k = 0; k < n_ent; k++)
                if (!(hypot(vals[k].x, vals[k].y) < DBL_EPSILON)) c++;
            counts[j] = c;
        } else {
            for (int a = 1; a < n_ent; a++) {                      // ascending columns (std::map order)
                const int64_t ck = cols[a];
                const double2 vk = vals[a];
                int b = a - 1;
                while (b >= 0 && cols[b] > ck) {
                    cols[b + 1] = cols[b];
                    vals[b + 1] = vals[b];
                    b--;
                }
                cols[b + 1] = ck;
                vals[b + 1] = vk;
            }
            int64_t o = row_ptr[j];
            for (int k = 0; k < n_ent; k++) {
                if (hypot(vals[k].x, vals[k].y) < DBL_EPSILON) continue;   // no zeros (:381-391)
                col_out[o] = (int)cols[k];
                val_out[o] = vals[k];
                o++;
            }
        }
    }
}

__global__ void narrow_ptr_kernel(const int64_t* in, int64_t n, int* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) out[i] = (int)in[i];
}

// ---- exact norm1: column sums of |v| accumulated in row order, as the reference's COO loop ----
__global__ void col_count_kernel(int64_t n_rows, const int* ptr, const int* col, int64_t* cnt) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    for (int k = ptr[r]; k < ptr[r + 1]; k++) atomicAdd((unsigned long long*)&cnt[col[k]], 1ull);
}

__global__ void col_scatter_kernel(int64_t n_rows, const int* ptr, const int* col, const double2* val,
                                   const int64_t* cptr, int64_t* fill, int* trow, double* tabs) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    for (int k = ptr[r]; k < ptr[r + 1]; k++) {
        const int c = col[k];
        const int64_t pos = cptr[c] + (int64_t)atomicAdd((unsigned long long*)&fill[c], 1ull);
        trow[pos] = (int)r;
        tabs[pos] = hypot(val[k].x, val[k].y);
    }
}

__global__ void col_sum_kernel(int64_t n_cols, const int64_t* cptr, int* trow, double* tabs, double* colsum) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cols) return;
    const int64_t a = cptr[c], b = cptr[c + 1];
    for (int64_t i = a + 1; i < b; i++) {                 // sort the column's entries by row
        const int rk = trow[i];
        const double vk = tabs[i];
        int64_t q = i - 1;
        while (q >= a && trow[q] > rk) {
            trow[q + 1] = trow[q];
            tabs[q + 1] = tabs[q];
            q--;
        }
        trow[q + 1] = rk;
        tabs[q + 1] = vk;
    }
    double s = 0.0;
    for (int64_t i = a; i < b; i++) s += tabs[i];
    colsum[c] = s;
}

__global__ void offdiag_count_kernel(int64_t n_rows, const int* ptr, const int* col, const double2* val,
                                     unsigned long long* off, double* diag) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    double d = 0.0;
    for (int k = ptr[r]; k < ptr[r + 1]; k++) {
        if (col[k] != r) atomicAdd(off, 1ull);
        else d += val[k].x;
    }
    diag[r] = d;
}

static int finalize_matrix(te_ctx* ctx, te_matrix* A) {
    hipStream_t s = ctx->stream;
    const int64_t n = A->n_rows;
    // norm1
    int64_t *cnt = nullptr, *cptr = nullptr, *fill = nullptr;
    int* trow = nullptr;
    double *tabs = nullptr, *colsum = nullptr;
    TE_HIP(ctx, hipMalloc(&cnt, sizeof(int64_t) * (A->n_cols + 1)));
    TE_HIP(ctx, hipMalloc(&cptr, sizeof(int64_t) * (A->n_cols + 1)));
    TE_HIP(ctx, hipMalloc(&fill, sizeof(int64_t) * (A->n_cols + 1)));
    TE_HIP(ctx, hipMalloc(&trow, sizeof(int) * std::max<int64_t>(A->nnz, 1)));
    TE_HIP(ctx, hipMalloc(&tabs, sizeof(double) * std::max<int64_t>(A->nnz, 1)));
    TE_HIP(ctx, hipMalloc(&colsum, sizeof(double) * A->n_cols));
    TE_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(int64_t) * (A->n_cols + 1), s));
    TE_HIP(ctx, hipMemsetAsync(fill, 0, sizeof(int64_t) * (A->n_cols + 1), s));
    hipLaunchKernelGGL(col_count_kernel, dim3(blocks_for(n)), dim3(256), 0, s, n, A->row_ptr, A->col, cnt);
    int rc = exclusive_scan(ctx, cnt, cptr, A->n_cols + 1);
    if (rc != TE_OK) return rc;
    hipLaunchKernelGGL(col_scatter_kernel, dim3(blocks_for(n)), dim3(256), 0, s, n, A->row_ptr, A->col, A->val, cptr,
                       fill, trow, tabs);
    hipLaunchKernelGGL(col_sum_kernel, dim3(blocks_for(A->n_cols)), dim3(256), 0, s, A->n_cols, cptr, trow, tabs,
                       colsum);
    std::vector<double> h(A->n_cols);
    TE_HIP(ctx, hipMemcpyAsync(h.data(), colsum, sizeof(double) * A->n_cols, hipMemcpyDeviceToHost, s));
    // diagonal form
    unsigned long long* offc = nullptr;
    double* diag = nullptr;
    TE_HIP(ctx, hipMalloc(&offc, sizeof(unsigned long long)));
    TE_HIP(ctx, hipMalloc(&diag, sizeof(double) * n));
    TE_HIP(ctx, hipMemsetAsync(offc, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(offdiag_count_kernel, dim3(blocks_for(n)), dim3(256), 0, s, n, A->row_ptr, A->col, A->val, offc,
                       diag);
    unsigned long long noff = 0;
    TE_HIP(ctx, hipMemcpyAsync(&noff, offc, sizeof(noff), hipMemcpyDeviceToHost, s));
    TE_HIP(ctx, hipStreamSynchronize(s));
    A->norm1 = h.empty() ? 0.0 : *std::max_element(h.begin(), h.end());
    A->norm1_valid = true;
    if (noff == 0 && A->n_rows == A->n_cols) {
        A->diagonal = true;
        A->diag_re = diag;
    } else {
        hipFree(diag);
    }
    hipFree(cnt);
    hipFree(cptr);
    hipFree(fill);
    hipFree(trow);
    hipFree(tabs);
    hipFree(colsum);
    hipFree(offc);
    return TE_OK;
}

}  // namespace te

using namespace te;

static int basis_finish(te_ctx* ctx, te_basis* B, te_basis** out) {
    *out = B;
    return TE_OK;
}

extern "C" int te_basis_create(te_ctx* ctx, int N, int K, int nbQudits, int capacity, int quant, te_basis** out) {
    if (!ctx || !out) return TE_ERR_ARG;
    *out = nullptr;
    if (K < 1 || K > kMaxModes || N < 0 || quant < 1 || nbQudits < 0 || capacity < 0)
        return fail(ctx, TE_ERR_ARG, "unsupported basis parameters (1 <= K <= %d, N >= 0, quant >= 1)", kMaxModes);
    hipSetDevice(ctx->device);
    te_basis* B = new te_basis();
    B->ctx = ctx;
    B->K = K;
    int* occ = nullptr;
    int rc = build_factor(ctx, B, N, K, nbQudits, capacity, quant, B->dev.f[0], &occ);
    if (rc != TE_OK) {
        te_basis_destroy(B);
        return rc;
    }
    B->occ = occ;
    B->D = B->dev.f[0].D;
    B->max_occ = N;
    B->dev.K = K;
    B->dev.K1 = K;
    B->dev.nf = 1;
    B->dev.D = B->D;
    B->dev.occ = occ;
    return basis_finish(ctx, B, out);
}

extern "C" int te_tensor_basis_create(te_ctx* ctx, int N1, int K1, int N2, int K2, int C2, te_basis** out) {
    if (!ctx || !out) return TE_ERR_ARG;
    *out = nullptr;
    if (K1 < 1 || K2 < 1 || K1 + K2 > kMaxModes || N1 < 0 || N2 < 0 || C2 < 0)
        return fail(ctx, TE_ERR_ARG, "unsupported tensor basis parameters");
    hipSetDevice(ctx->device);
    te_basis* B = new te_basis();
    B->ctx = ctx;
    B->K = K1 + K2;
    int *o1 = nullptr, *o2 = nullptr;
    int rc = build_factor(ctx, B, N1, K1, 0, 0, 1, B->dev.f[0], &o1);
    if (rc == TE_OK) rc = build_factor(ctx, B, N2, K2, K2, C2, 1, B->dev.f[1], &o2);
    if (rc != TE_OK) {
        if (o1) hipFree(o1);
        if (o2) hipFree(o2);
        te_basis_destroy(B);
        return rc;
    }
    const int64_t n1 = B->dev.f[0].D, n2 = B->dev.f[1].D;
    B->D = n1 * n2;                          // numberElements = length1 * length2 (:143)
    int* occ = nullptr;
    if (hipMalloc(&occ, sizeof(int) * std::max<int64_t>(B->D, 1) * B->K) != hipSuccess) {
        hipFree(o1);
        hipFree(o2);
        te_basis_destroy(B);
        return fail(ctx, TE_ERR_ALLOC, "tensor basis allocation failed");
    }
    if (B->D > 0)
        hipLaunchKernelGGL(tensor_kernel, dim3(blocks_for(B->D)), dim3(256), 0, ctx->stream, o1, n1, K1, o2, n2, K2,
                           occ);
    hipStreamSynchronize(ctx->stream);
    hipFree(o1);
    hipFree(o2);
    B->occ = occ;
    B->max_occ = std::max(N1, N2);
    B->dev.K = B->K;
    B->dev.K1 = K1;
    B->dev.nf = 2;
    B->dev.D = B->D;
    B->dev.occ = occ;
    return basis_finish(ctx, B, out);
}

extern "C" void te_basis_destroy(te_basis* B) {
    if (!B) return;
    if (B->occ) hipFree(B->occ);
    if (B->lex_perm) hipFree(B->lex_perm);
    for (void* p : B->allocations) hipFree(p);
    delete B;
}

extern "C" int te_basis_info(const te_basis* B, size_t* n_elements, int* n_modes) {
    if (!B) return TE_ERR_ARG;
    if (n_elements) *n_elements = (size_t)B->D;
    if (n_modes) *n_modes = B->K;
    return TE_OK;
}

extern "C" int te_basis_states(const te_basis* B, int* occ) {
    if (!B || !occ) return TE_ERR_ARG;
    te_ctx* ctx = B->ctx;
    hipSetDevice(ctx->device);
    if (B->D > 0) TE_HIP(ctx, hipMemcpy(occ, B->occ, sizeof(int) * B->D * B->K, hipMemcpyDeviceToHost));
    return TE_OK;
}

extern "C" int te_basis_find(const te_basis* B, const int* occ, size_t count, int64_t* index_out) {
    if (!B || (count && (!occ || !index_out))) return TE_ERR_ARG;
    if (count == 0) return TE_OK;
    te_ctx* ctx = B->ctx;
    hipSetDevice(ctx->device);
    int* d_occ = nullptr;
    int64_t* d_out = nullptr;
    TE_HIP(ctx, hipMalloc(&d_occ, sizeof(int) * count * B->K));
    TE_HIP(ctx, hipMalloc(&d_out, sizeof(int64_t) * count));
    TE_HIP(ctx, hipMemcpyAsync(d_occ, occ, sizeof(int) * count * B->K, hipMemcpyHostToDevice, ctx->stream));
    if (B->D > 0) {
        hipLaunchKernelGGL(find_kernel, dim3(blocks_for((int64_t)count)), dim3(256), 0, ctx->stream, B->dev, d_occ,
                           (int64_t)count, d_out);
    } else {
        TE_HIP(ctx, hipMemsetAsync(d_out, 0xff, sizeof(int64_t) * count, ctx->stream));
    }
    TE_HIP(ctx, hipMemcpyAsync(index_out, d_out, sizeof(int64_t) * count, hipMemcpyDeviceToHost, ctx->stream));
    TE_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(d_occ);
    hipFree(d_out);
    return TE_OK;
}

namespace te {

// sig[p]: the mode at significance position p (0 = most significant).
struct ModeOrder {
    int sig[kMaxModes];
};

__global__ void lex_key_kernel(const int* __restrict__ occ, int64_t D, int K, unsigned long long radix,
                               ModeOrder mo, unsigned long long* keys, int* idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    unsigned long long k = 0;
    for (int p = 0; p < K; p++) k = k * radix + (unsigned long long)occ[i * K + mo.sig[p]];
    keys[i] = k;
    idx[i] = (int)i;
}

// Significance order of the modes for the locality order: a Cuthill-McKee numbering of the graph
// whose vertices are the modes and whose edges join the modes of one term, from the start mode
// that gives the smallest bandwidth (ties: the lowest start). A transition that changes modes at
// significance positions p < q moves a state's rank by about the number of states sharing its
// first p modes, and every row of a 64-row slice shares its leading modes, so with a small
// bandwidth most transitions change only leading modes that a whole slice shares: their column
// offset and value are the same on every row of the slice. On a ring of K modes this gives
// 0, 1, K-1, 2, K-2, ... (bandwidth 2; mode 0 first and last would put the closing bond's two
// modes at the two ends). Without two-mode terms: the modes in index order.
static ModeOrder mode_significance(int K, const std::vector<int>& off, const std::vector<int>& modes) {
    ModeOrder mo{};
    if (K > kMaxModes) return mo;                        // the key cannot hold such a basis anyway
    for (int m = 0; m < K; m++) mo.sig[m] = m;
    std::vector<std::vector<char>> adj(K, std::vector<char>(K, 0));
    bool any = false;
    for (size_t t = 0; t + 1 < off.size(); t++)
        for (int a = off[t]; a < off[t + 1]; a++)
            for (int b = off[t]; b < off[t + 1]; b++)
                if (modes[a] != modes[b]) adj[modes[a]][modes[b]] = 1, any = true;
    if (!any) return mo;
    std::vector<int> deg(K, 0);
    for (int u = 0; u < K; u++)
        for (int v = 0; v < K; v++) deg[u] += adj[u][v];
    int best_bw = K + 1;
    for (int start = 0; start < K; start++) {
        std::vector<int> order, pos(K, -1);
        order.reserve(K);
        // Cuthill-McKee from `start`; a disconnected remainder restarts at its lowest mode
        for (int seed = start, done = 0; done < K;) {
            if (pos[seed] < 0) {
                pos[seed] = (int)order.size();
                order.push_back(seed);
                for (size_t h = order.size() - 1; h < order.size(); h++) {
                    std::vector<int> nb;
                    for (int v = 0; v < K; v++)
                        if (adj[order[h]][v] && pos[v] < 0) nb.push_back(v);
                    std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
                    for (int v : nb) {
                        pos[v] = (int)order.size();
                        order.push_back(v);
                    }
                }
            }
            done = (int)order.size();
            seed = 0;
            while (seed < K && pos[seed] >= 0) seed++;
            if (seed == K) break;
        }
        int bw = 0;
        for (int u = 0; u < K; u++)
            for (int v = 0; v < K; v++)
                if (adj[u][v]) bw = std::max(bw, std::abs(pos[u] - pos[v]));
        if (bw < best_bw) {
            best_bw = bw;
            for (int p = 0; p < K; p++) mo.sig[p] = order[p];
        }
    }
    return mo;
}

// Lexicographic order of the basis states (modes in the significance order `mo`, mixed radix
// max_occ + 1), as new -> old. Neighbouring states that differ in the trailing modes get
// neighbouring indices, so a product's gathers from x stay within a few thousand entries for the
// transitions among those modes. Computed once per basis, by the first operator with two-mode
// terms that is built on it. Null (identity) when the key does not fit 64 bits or
// TE_SPMV_ORDER=0; TE_SPMV_ORDER=lex keeps the modes in index order.
static int basis_locality_order(te_ctx* ctx, const te_basis* B, ModeOrder mo, const int** out) {
    *out = nullptr;
    const char* e = std::getenv("TE_SPMV_ORDER");
    if (e && e[0] == '0') return TE_OK;
    if (B->lex_done) {
        *out = B->lex_perm;
        return TE_OK;
    }
    B->lex_done = true;
    if (B->D < 2 || B->D > INT32_MAX) return TE_OK;
    const unsigned long long radix = (unsigned long long)B->max_occ + 1;
    unsigned __int128 span = 1;
    for (int m = 0; m < B->K; m++) {
        span *= radix;
        if (span > ((unsigned __int128)1 << 64)) return TE_OK;
    }
    int bits = 0;
    while (bits < 64 && ((unsigned __int128)1 << bits) < span) bits++;
    const int64_t D = B->D;
    hipStream_t s = ctx->stream;
    unsigned long long *k_in = nullptr, *k_out = nullptr;
    int *i_in = nullptr, *perm = nullptr;
    void* tmp = nullptr;
    size_t bytes = 0;
    int rc = TE_OK;
    if (hipMalloc(&k_in, sizeof(unsigned long long) * D) != hipSuccess ||
        hipMalloc(&k_out, sizeof(unsigned long long) * D) != hipSuccess ||
        hipMalloc(&i_in, sizeof(int) * D) != hipSuccess || hipMalloc(&perm, sizeof(int) * D) != hipSuccess)
        rc = fail(ctx, TE_ERR_ALLOC, "locality order allocation");
    if (rc == TE_OK) {
        hipLaunchKernelGGL(lex_key_kernel, dim3(blocks_for(D)), dim3(256), 0, s, B->occ, D, B->K, radix, mo, k_in,
                           i_in);
        if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k_in, k_out, i_in, perm, (int)D, 0, bits, s) !=
                hipSuccess ||
            hipMalloc(&tmp, std::max<size_t>(bytes, 1)) != hipSuccess)
            rc = fail(ctx, TE_ERR_HIP, "locality order sort setup");
    }
    if (rc == TE_OK && (hipcub::DeviceRadixSort::SortPairs(tmp, bytes, k_in, k_out, i_in, perm, (int)D, 0, bits, s) !=
                            hipSuccess ||
                        hipStreamSynchronize(s) != hipSuccess))
        rc = fail(ctx, TE_ERR_HIP, "locality order sort");
    void* frees[] = {k_in, k_out, i_in, tmp};
    for (void* p : frees)
        if (p) hipFree(p);
    if (rc != TE_OK) {
        if (perm) hipFree(perm);
        return rc;
    }
    B->lex_perm = perm;
    *out = perm;
    return TE_OK;
}

}  // namespace te

extern "C" int te_operator_build(te_ctx* ctx, const te_basis* B, const te_op_term* terms, int n_terms, te_matrix** out) {
    if (!ctx || !B || !out || n_terms < 0 || (n_terms > 0 && !terms)) return TE_ERR_ARG;
    *out = nullptr;
    if (B->D == 0) return fail(ctx, TE_ERR_FATAL, "Empty matrices are not supported.");   // matrixDataTypes.cpp:110-113
    hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    // flatten the terms; validate modes (creationOperator exits with 14 otherwise, :412-416)
    std::vector<int> off(1, 0), ops, modes, caps;
    std::vector<double2> coef;
    int max_ops = 0;
    for (int t = 0; t < n_terms; t++) {
        const te_op_term& T = terms[t];
        if (T.n_ops < 0 || (T.n_ops > 0 && (!T.operations || !T.modes || !T.capacities)))
            return fail(ctx, TE_ERR_ARG, "malformed op term %d", t);
        for (int i = 0; i < T.n_ops; i++) {
            if (T.modes[i] < 0 || T.modes[i] >= B->K)
                return fail(ctx, TE_ERR_FATAL, "creationOperator: mode does not exist in vector");
            ops.push_back(T.operations[i]);
            modes.push_back(T.modes[i]);
            caps.push_back(T.capacities[i]);
        }
        off.push_back((int)ops.size());
        coef.push_back(make_double2(T.coef_re, T.coef_im));
        max_ops = std::max(max_ops, T.n_ops);
    }
    // largest occupation reachable: the basis maximum plus one per operator
    const int sqrt_n = B->max_occ + max_ops + 2;
    std::vector<double> sqrt_tab(sqrt_n);
    for (int k = 0; k < sqrt_n; k++) sqrt_tab[k] = std::sqrt((double)k);   // correctly rounded (IEEE)

    auto up = [&](const void* h, size_t bytes, void** d) -> int {
        TE_HIP(ctx, hipMalloc(d, std::max<size_t>(bytes, 16)));
        if (bytes) TE_HIP(ctx, hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
        return TE_OK;
    };
    void *d_off = nullptr, *d_ops = nullptr, *d_modes = nullptr, *d_caps = nullptr, *d_coef = nullptr, *d_sqrt = nullptr;
    int rc = TE_OK;
    if (rc == TE_OK) rc = up(off.data(), sizeof(int) * off.size(), &d_off);
    if (rc == TE_OK) rc = up(ops.data(), sizeof(int) * ops.size(), &d_ops);
    if (rc == TE_OK) rc = up(modes.data(), sizeof(int) * modes.size(), &d_modes);
    if (rc == TE_OK) rc = up(caps.data(), sizeof(int) * caps.size(), &d_caps);
    if (rc == TE_OK) rc = up(coef.data(), sizeof(double2) * coef.size(), &d_coef);
    if (rc == TE_OK) rc = up(sqrt_tab.data(), sizeof(double) * sqrt_tab.size(), &d_sqrt);
    TermsDev T{};
    T.n_terms = n_terms;
    T.max_entries = std::max(n_terms, 1);
    T.term_off = (const int*)d_off;
    T.ops = (const int*)d_ops;
    T.modes = (const int*)d_modes;
    T.caps = (const int*)d_caps;
    T.coef = (const double2*)d_coef;
    T.sqrt_tab = (const double*)d_sqrt;
    T.sqrt_n = sqrt_n;

    const size_t per = ((size_t)T.max_entries * 24 + (size_t)2 * B->K * 4 + 15) / 16 * 16;
    const size_t lds_cap = (size_t)ctx->max_lds * 3 / 4;
    int bs = 256;
    while (bs > 64 && per * bs > lds_cap) bs >>= 1;
    const bool use_lds = per * bs <= lds_cap;
    const int grid = std::max(1, std::min<int>((int)blocks_for(B->D, bs), ctx->num_cus * (use_lds ? 8 : 2)));
    if (use_lds) {
        hipFuncSetAttribute((const void*)operator_rows_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(per * bs));
        hipFuncSetAttribute((const void*)operator_rows_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(per * bs));
    }
    char* ws = nullptr;
    if (rc == TE_OK && !use_lds) {
        if (hipMalloc(&ws, per * (size_t)grid * bs) != hipSuccess) rc = fail(ctx, TE_ERR_ALLOC, "workspace");
    }
    const size_t shmem = use_lds ? per * bs : 0;
    int64_t *counts = nullptr, *rp64 = nullptr;
    te_matrix* A = nullptr;
    if (rc == TE_OK) {
        if (hipMalloc(&counts, sizeof(int64_t) * (B->D + 1)) != hipSuccess ||
            hipMalloc(&rp64, sizeof(int64_t) * (B->D + 1)) != hipSuccess)
            rc = fail(ctx, TE_ERR_ALLOC, "row counts");
    }
    int64_t nnz = 0;
    if (rc == TE_OK) {
        hipMemsetAsync(counts + B->D, 0, sizeof(int64_t), s);
        hipLaunchKernelGGL(operator_rows_kernel<false>, dim3(grid), dim3(bs), shmem, s, B->dev, T, ws, counts,
                           (const int64_t*)nullptr, (int*)nullptr, (double2*)nullptr);
        if (hipGetLastError() != hipSuccess) rc = fail(ctx, TE_ERR_HIP, "operator count kernel launch failed");
    }
    if (rc == TE_OK) rc = exclusive_scan(ctx, counts, rp64, B->D + 1);
    if (rc == TE_OK) {
        if (hipMemcpy(&nnz, rp64 + B->D, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(ctx, TE_ERR_HIP, "nnz copy");
        else if (nnz >= (int64_t)INT32_MAX)
            rc = fail(ctx, TE_ERR_UNSUPPORTED, "operator has %lld >= 2^31 non-zeros", (long long)nnz);
    }
    if (rc == TE_OK) {
        A = new te_matrix();
        A->ctx = ctx;
        A->n_rows = A->n_cols = B->D;
        A->nnz = nnz;
        if (hipMalloc(&A->row_ptr, sizeof(int) * (B->D + 1)) != hipSuccess ||
            hipMalloc(&A->col, sizeof(int) * std::max<int64_t>(nnz, 1)) != hipSuccess ||
            hipMalloc(&A->val, sizeof(double2) * std::max<int64_t>(nnz, 1)) != hipSuccess)
            rc = fail(ctx, TE_ERR_ALLOC, "CSR allocation (%lld non-zeros)", (long long)nnz);
    }
    if (rc == TE_OK) {
        hipLaunchKernelGGL(narrow_ptr_kernel, dim3(blocks_for(B->D + 1)), dim3(256), 0, s, rp64, B->D, A->row_ptr);
        hipLaunchKernelGGL(operator_rows_kernel<true>, dim3(grid), dim3(bs), shmem, s, B->dev, T, ws, (int64_t*)nullptr,
                           rp64, A->col, A->val);
        if (hipGetLastError() != hipSuccess) rc = fail(ctx, TE_ERR_HIP, "operator fill kernel launch failed");
        else if (hipStreamSynchronize(s) != hipSuccess) rc = fail(ctx, TE_ERR_HIP, "operator build failed");
    }
    if (rc == TE_OK) rc = finalize_matrix(ctx, A);
    const int* order = nullptr;
    if (rc == TE_OK && !A->diagonal) {
        ModeOrder mo = mode_significance(B->K, off, modes);
        const char* e = std::getenv("TE_SPMV_ORDER");
        if (e && std::strcmp(e, "lex") == 0)
            for (int m = 0; m < B->K; m++) mo.sig[m] = m;
        rc = basis_locality_order(ctx, B, mo, &order);
    }
    if (rc == TE_OK) rc = prepare_spmv(ctx, A, order);
    void* frees[] = {d_off, d_ops, d_modes, d_caps, d_coef, d_sqrt, ws, counts, rp64};
    for (void* p : frees)
        if (p) hipFree(p);
    if (rc != TE_OK) {
        if (A) te_matrix_destroy(A);
        return rc;
    }
    *out = A;
    return TE_OK;
}
