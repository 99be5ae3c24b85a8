// Complex128 CSR SpMV and the fused Lanczos kernels for gfx950.
//
// Replaces smatrix::spMV = mkl_sparse_z_mv (core/matrixDataTypes.cpp:169-190) and the
// BLAS-1 chain of arnoldiAlgorithm (core/krylovTimeEvolver.cpp:523-571).
//
// Layout: CSR with int32 row pointers and column indices, complex128 values as double2
// (16-byte loads). A group of LPR lanes of a wavefront shares one row: lane l reads entries
// start+l, start+l+LPR, ..., so neighbouring groups stream neighbouring value/index ranges
// and a wave issues wide, mostly contiguous loads. Partial row sums are combined with
// xor-shuffles inside the group. Blocks own contiguous row ranges (xcd_block_id keeps an
// XCD's rows contiguous so the gather of x reuses its L2).
//
// Lanczos in unnormalised form: the Krylov vectors are stored as u_j with v_j = s_j u_j,
// s_0 = 1, s_{j+1} = 1/beta_j, so no separate normalisation pass exists:
//   apply  : w = (alpha s_j) A u_j + (beta_{j-1} s_{j-1}) u_{j-1};  a_j = s_j <u_j, w>
//   project: w' = w - (a_j s_j) u_j;  beta_j = |w'|;  u_{j+1} = w' (stored in place)
// which is the reference's recurrence (:530-565) with H[j][j] = a_j, H[j+1][j] = beta_j.
#include "te_internal.h"
#include "te_krylov.h"
#include "te_reduce.h"
#include "te_sell.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace te {

// ---------------------------------------------------------------------------------------
// Row-range partition: block `lb` owns rows [lb*chunk, min((lb+1)*chunk, n)).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void block_rows(int64_t n, int64_t& lo, int64_t& hi) {
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int lb = xcd_block_id();
    lo = (int64_t)lb * chunk;
    hi = lo + chunk < n ? lo + chunk : n;
}

// Sum over the row's entries of val[k] * x[col[k]] by the LPR lanes of a group.
template <int LPR>
__device__ __forceinline__ double2 row_dot(int row, const int* __restrict__ ptr, const int* __restrict__ col,
                                           const double2* __restrict__ val, const double2* __restrict__ x,
                                           int lane) {
    const int start = ptr[row], end = ptr[row + 1];
    double sr = 0.0, si = 0.0;
    for (int k = start + lane; k < end; k += LPR) {
        const double2 a = val[k];
        const double2 b = x[col[k]];
        sr += a.x * b.x - a.y * b.y;
        si += a.x * b.y + a.y * b.x;
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off, LPR);
        si += __shfl_xor(si, off, LPR);
    }
    return make_double2(sr, si);
}

template <int LPR>
__global__ __launch_bounds__(kBlock) void spmv_kernel(int64_t n, const int* __restrict__ ptr,
                                                      const int* __restrict__ col, const double2* __restrict__ val,
                                                      const double2* __restrict__ x, double2* __restrict__ y,
                                                      double2 alpha) {
    constexpr int RPB = kBlock / LPR;
    int64_t lo, hi;
    block_rows(n, lo, hi);
    const int lane = threadIdx.x % LPR;
    for (int64_t base = lo; base < hi; base += RPB) {
        const int64_t row = base + threadIdx.x / LPR;
        if (row < hi) {
            double2 s = row_dot<LPR>((int)row, ptr, col, val, x, lane);
            if (lane == 0) y[row] = cmul(alpha, s);
        }
    }
}

template <int LPR>
__global__ __launch_bounds__(kBlock) void lanczos_apply_kernel(LanczosArgs a) {
    KrylovScalars* ks = a.ks;
    if (ks->broken) return;
    constexpr int RPB = kBlock / LPR;
    __shared__ double2 sh[kBlock / 64];
    const int j = a.j;
    const double sj = ks->s[j];
    const double2 c1 = make_double2(a.alpha.x * sj, a.alpha.y * sj);
    const double c2 = j > 0 ? ks->beta[j - 1] * ks->s[j - 1] : 0.0;
    const double2* __restrict__ uj = a.U + (size_t)j * a.ld;
    const double2* __restrict__ up = j > 0 ? a.U + (size_t)(j - 1) * a.ld : nullptr;
    double2* __restrict__ w = a.U + (size_t)(j + 1) * a.ld;
    int64_t lo, hi;
    block_rows(a.n, lo, hi);
    const int lane = threadIdx.x % LPR;
    double2 dot = make_double2(0.0, 0.0);
    for (int64_t base = lo; base < hi; base += RPB) {
        const int64_t row = base + threadIdx.x / LPR;
        if (row < hi) {
            double2 s = row_dot<LPR>((int)row, a.ptr, a.col, a.val, uj, lane);
            if (lane == 0) {
                double2 wr = cmul(c1, s);
                if (j > 0) {
                    const double2 p = up[row];
                    wr.x += c2 * p.x;
                    wr.y += c2 * p.y;
                }
                w[row] = wr;
                const double2 u = uj[row];
                dot.x += u.x * wr.x + u.y * wr.y;
                dot.y += u.x * wr.y - u.y * wr.x;
            }
        }
    }
    double2 bs = block_sum2<kBlock>(dot, sh);
    if (threadIdx.x == 0) a.partials[blockIdx.x] = bs;
    if (last_block_arrive(a.ticket)) {
        double2 tot = reduce_partials<kBlock>(a.partials, gridDim.x, 1, sh);
        if (threadIdx.x == 0) {
            ks->a[j] = make_double2(sj * tot.x, sj * tot.y);
            *a.ticket = 0;
        }
    }
}

// Rows [lo, hi) of the projection w' = w - c3 u_j (stored unless j + 1 == m); returns their
// contribution to |w'|^2.
__device__ __forceinline__ double project_rows(const LanczosArgs& a, int64_t lo, int64_t hi) {
    const KrylovScalars* ks = a.ks;
    const int j = a.j;
    const double sj = ks->s[j];
    const double2 aj = ks->a[j];
    const double2 c3 = make_double2(aj.x * sj, aj.y * sj);
    const double2* __restrict__ uj = a.U + (size_t)j * a.ld;
    double2* __restrict__ w = a.U + (size_t)(j + 1) * a.ld;
    const bool store = j + 1 < a.m;
    double nrm = 0.0;
    int64_t row = lo + threadIdx.x;
    // four rows per thread in flight: all eight loads are issued before any is used. Both are
    // read non-temporally (3 % off the launch), and u_{j+1} is stored non-temporally too (on the
    // L = N = 14 ring the next apply took 610.3-610.6 us against 613 us after ordinary stores,
    // time-to-T 0.4351-0.4355 s against 0.4366-0.4370 s).
    for (; row + 3 * kBlock < hi; row += 4 * kBlock) {
        double2 u[4], wr[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            u[q] = load_stream(&uj[row + q * kBlock]);
            wr[q] = load_stream(&w[row + q * kBlock]);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            wr[q].x -= c3.x * u[q].x - c3.y * u[q].y;
            wr[q].y -= c3.x * u[q].y + c3.y * u[q].x;
            if (store) store_stream(&w[row + q * kBlock], wr[q]);
            nrm += wr[q].x * wr[q].x + wr[q].y * wr[q].y;
        }
    }
    for (; row < hi; row += kBlock) {
        const double2 u = load_stream(&uj[row]);
        double2 wr = load_stream(&w[row]);
        wr.x -= c3.x * u.x - c3.y * u.y;
        wr.y -= c3.x * u.y + c3.y * u.x;
        if (store) store_stream(&w[row], wr);
        nrm += wr.x * wr.x + wr.y * wr.y;
    }
    return nrm;
}

// The finish of a projection from beta^2 = tot.x.
__device__ __forceinline__ void project_finish(const LanczosArgs& a, double2 tot) {
    KrylovScalars* ks = a.ks;
    const int j = a.j;
    const double beta = sqrt(tot.x);
    if (beta < a.tolRate) {            // lucky breakdown (core/krylovTimeEvolver.cpp:551-555)
        ks->broken = 1;
        ks->m_hbd = j + 1;
        ks->h = beta;
    } else if (j + 1 == a.m) {
        ks->h = beta;                  // :566-567
    } else {
        ks->beta[j] = beta;            // :558-565
        ks->s[j + 1] = 1.0 / beta;
    }
}

// The projection's rows in chunks of kProjectChunk, block b taking chunks nch-1-b,
// nch-1-b-grid, ...: all blocks sweep one front from the last rows to the first. The apply's
// front ends on the last rows, so the projection starts on the w rows still held in the MALL,
// and ends on the u_{j+1} rows the next apply gathers first (5 % off the projection, 1 % off the
// apply on the L = N = 14 ring, against a contiguous range per block). Returns this block's
// contribution to |w'|^2 (the sum over its chunks in that order).
constexpr int64_t kProjectChunk = 4 * kBlock;

__device__ __forceinline__ double project_front(const LanczosArgs& a) {
    const int64_t nch = (a.n + kProjectChunk - 1) / kProjectChunk;
    double nrm = 0.0;
    for (int64_t k = blockIdx.x; k < nch; k += gridDim.x) {
        const int64_t lo = (nch - 1 - k) * kProjectChunk;
        nrm += project_rows(a, lo, lo + kProjectChunk < a.n ? lo + kProjectChunk : a.n);
    }
    return nrm;
}

// Ends at each block's partial of |w'|^2, as the sliced-ELL apply does (lanczos_finish_kernel).
__global__ __launch_bounds__(kBlock) void lanczos_project_kernel(LanczosArgs a) {
    if (a.ks->broken) return;
    __shared__ double2 sh[kBlock / 64];
    const double nrm = project_front(a);
    double2 bs = block_sum2<kBlock>(make_double2(nrm, 0.0), sh);
    if (threadIdx.x == 0) a.partials[blockIdx.x] = bs;
}

// The projections of two recurrences in one launch (a and b: own partials); one that broke down
// is skipped. Each gets the row ranges and partials of its own launch.
__global__ __launch_bounds__(kBlock) void lanczos_project2_kernel(LanczosArgs a, LanczosArgs b) {
    const bool la = !a.ks->broken, lb = !b.ks->broken;
    if (!la && !lb) return;
    __shared__ double2 sh[kBlock / 64];
    const double n0 = la ? project_front(a) : 0.0;
    const double n1 = lb ? project_front(b) : 0.0;
    const double2 bs0 = block_sum2<kBlock>(make_double2(n0, 0.0), sh);
    const double2 bs1 = block_sum2<kBlock>(make_double2(n1, 0.0), sh);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = bs0;
        b.partials[blockIdx.x] = bs1;
    }
}

// ---------------------------------------------------------------------------------------
// Row-block ("CSR-stream") kernels. Each block walks a contiguous range of row blocks; for one
// row block every thread issues all of its value, index and x loads (kStreamCap / kBlock of
// each) before using any, the products go to LDS, and thread t then sums rows r0 + t, ... in
// entry order.
// ---------------------------------------------------------------------------------------
template <class Epi>
__device__ __forceinline__ void stream_rows(const int* __restrict__ rb, int nrb, const int* __restrict__ ptr,
                                            const int* __restrict__ col, const double2* __restrict__ val,
                                            const double2* __restrict__ x, double2* prod, Epi& epi,
                                            const double2* __restrict__ acc_in) {
    int b0, b1, bstride;
    xcd_sweep(nrb, b0, b1, bstride);
    for (int b = b0; b < b1; b += bstride) stream_row_block(rb[b], rb[b + 1], ptr, col, val, x, prod, epi, acc_in);
}

struct SpmvEpi {
    double2 alpha;
    double2* y;
    __device__ __forceinline__ void row(int r, double2 s) { y[r] = cmul(alpha, s); }
};

__global__ __launch_bounds__(kBlock) void spmv_stream_kernel(const int* __restrict__ rb, int nrb,
                                                             const int* __restrict__ ptr, const int* __restrict__ col,
                                                             const double2* __restrict__ val,
                                                             const double2* __restrict__ x, double2* __restrict__ y,
                                                             double2 alpha, const double2* __restrict__ acc_in) {
    __shared__ double2 prod[kStreamCap];
    SpmvEpi epi{alpha, y};
    stream_rows(rb, nrb, ptr, col, val, x, prod, epi, acc_in);
}

struct PartialEpi {
    double2* acc;
    __device__ __forceinline__ void row(int r, double2 s) { store_stream(&acc[r], s); }
};

// One column slice of a split SpMV: acc[r] (+)= row r of the slice times x, unscaled.
__global__ __launch_bounds__(kBlock) void spmv_part_kernel(const int* __restrict__ rb, int nrb,
                                                           const int* __restrict__ ptr, const int* __restrict__ col,
                                                           const double2* __restrict__ val,
                                                           const double2* __restrict__ x, const double2* acc_in,
                                                           double2* acc, const int* broken) {
    if (broken && *broken) return;
    __shared__ double2 prod[kStreamCap];
    PartialEpi epi{acc};
    stream_rows(rb, nrb, ptr, col, val, x, prod, epi, acc_in);
}

struct ApplyEpi {
    double2 c1;
    double c2;
    const double2* up;
    const double2* uj;
    double2* w;
    double2 dot;
    __device__ __forceinline__ void row(int r, double2 s) {
        double2 wr = cmul(c1, s);
        if (up) {
            const double2 p = load_stream(&up[r]);
            wr.x += c2 * p.x;
            wr.y += c2 * p.y;
        }
        store_stream(&w[r], wr);
        const double2 u = uj[r];
        dot.x += u.x * wr.x + u.y * wr.y;
        dot.y += u.x * wr.y - u.y * wr.x;
    }
};

__global__ __launch_bounds__(kBlock) void lanczos_apply_stream_kernel(LanczosArgs a) {
    KrylovScalars* ks = a.ks;
    if (ks->broken) return;
    __shared__ double2 prod[kStreamCap];
    __shared__ double2 sh[kBlock / 64];
    const int j = a.j;
    const double sj = ks->s[j];
    ApplyEpi epi;
    epi.c1 = make_double2(a.alpha.x * sj, a.alpha.y * sj);
    epi.c2 = j > 0 ? ks->beta[j - 1] * ks->s[j - 1] : 0.0;
    epi.up = j > 0 ? a.U + (size_t)(j - 1) * a.ld : nullptr;
    epi.uj = a.U + (size_t)j * a.ld;
    epi.w = a.U + (size_t)(j + 1) * a.ld;
    epi.dot = make_double2(0.0, 0.0);
    stream_rows(a.row_block, a.n_row_blocks, a.ptr, a.col, a.val, epi.uj, prod, epi, a.acc_in);
    double2 bs = block_sum2<kBlock>(epi.dot, sh);
    if (threadIdx.x == 0) a.partials[blockIdx.x] = bs;
    if (last_block_arrive(a.ticket)) {
        double2 tot = reduce_partials<kBlock>(a.partials, gridDim.x, 1, sh);
        if (threadIdx.x == 0) {
            ks->a[j] = make_double2(sj * tot.x, sj * tot.y);
            *a.ticket = 0;
        }
    }
}

// ---------------------------------------------------------------------------------------
// Value-dictionary stream kernels. Same row blocks as above (chunk kVdChunk). Row r is
//   y_r = diag_r x_r + sum_k vtab[vidx_k] x[col_k]   (diagonal first, then column order)
// with the table staged in LDS once per block.
//
// Software pipeline over a block's row blocks: when the iteration for row block b starts, its
// columns and value indices are in registers and the bounds of the next one are known. The
// iteration issues b's gathers from x and its row loads (row pointers, diagonal, x_r and the
// epilogue's stream), the next block's column loads and the bounds of the one after, all before
// waiting on any of them, so a row block costs about one memory round trip instead of four.
// ---------------------------------------------------------------------------------------
constexpr int kVdNI = kVdCap / kBlock;
constexpr int kVdLds = kVdCap + kVdMaxValues;      // double2 entries: products, then the table

// Branch-free: slots past the block's end load entry k0 (kVdPad zero entries follow the last
// block, so k0 is in bounds also for an empty block at the end); their products are not stored.
__device__ __forceinline__ void vd_load_cols(const VdView& m, int k0, int nk, int* c, unsigned* vi) {
#pragma unroll
    for (int i = 0; i < kVdNI; i++) {
        const int k = threadIdx.x + i * kBlock;
        const int kk = k0 + (k < nk ? k : 0);
        c[i] = __builtin_nontemporal_load(&m.col[kk]);
        vi[i] = __builtin_nontemporal_load(&m.vidx[kk]);
    }
}

template <class Epi>
__device__ __forceinline__ void stream_rows_vd(const VdView& m, const double2* __restrict__ x, double2* lds, Epi& epi) {
    double2* prod = lds;
    double2* vt = lds + kVdCap;
    const int t = threadIdx.x;
    int b, end, stride;
    xcd_sweep(m.n_row_blocks, b, end, stride);
    int2 lo = make_int2(0, 0), hi = lo, nlo = lo, nhi = lo;
    int c[kVdNI];
    unsigned vi[kVdNI];
    if (b < end) {
        lo = m.blocks[b];
        hi = m.blocks[b + 1];
    }
    vd_load_cols(m, lo.y, hi.y - lo.y, c, vi);
    if (b + stride < end) {
        nlo = m.blocks[b + stride];
        nhi = m.blocks[b + stride + 1];
    }
    for (int i = t; i < m.n_values; i += kBlock) vt[i] = m.vtab[i];
    __syncthreads();
    for (; b < end; b += stride) {
        const int k0 = lo.y, nk = hi.y - lo.y;
        const int r0 = lo.x, nrows = hi.x - lo.x;
        double2 xv[kVdNI];
#pragma unroll
        for (int i = 0; i < kVdNI; i++) xv[i] = x[c[i]];
        // this thread's first row
        const int r = r0 + t;
        int pa = 0, pe = 0;
        double2 d = make_double2(0.0, 0.0), xr = d;
        typename Epi::Pre pre{};
        if (t < nrows) {
            pa = m.ptr[r] - k0;
            pe = m.ptr[r + 1] - k0;
            d = load_stream(&m.diag[r]);
            xr = x[r];
            pre = epi.load(r);
        }
        // the next row block's columns, the bounds of the one after it
        const int nb = b + stride;
        int cn[kVdNI];
        unsigned vin[kVdNI];
        vd_load_cols(m, nlo.y, nb < end ? nhi.y - nlo.y : 0, cn, vin);
        int2 nnlo = nlo, nnhi = nhi;
        if (nb + stride < end) {
            nnlo = m.blocks[nb + stride];
            nnhi = m.blocks[nb + stride + 1];
        }
#pragma unroll
        for (int i = 0; i < kVdNI; i++) {
            const int k = t + i * kBlock;
            if (k < nk) prod[k] = cmul(vt[vi[i]], xv[i]);
        }
        lds_barrier();
        if (t < nrows) {
            double sr = d.x * xr.x - d.y * xr.y, si = d.x * xr.y + d.y * xr.x;
            int k = pa;
            for (; k + 4 <= pe; k += 4) {        // four LDS reads in flight, sums in entry order
                const double2 p0 = prod[k], p1 = prod[k + 1], p2 = prod[k + 2], p3 = prod[k + 3];
                sr += p0.x;
                si += p0.y;
                sr += p1.x;
                si += p1.y;
                sr += p2.x;
                si += p2.y;
                sr += p3.x;
                si += p3.y;
            }
            for (; k < pe; k++) {
                sr += prod[k].x;
                si += prod[k].y;
            }
            epi.row(r, make_double2(sr, si), xr, pre);
        }
        for (int rr = r + kBlock; rr < r0 + nrows; rr += kBlock) {     // row blocks of many short rows
            const int a = m.ptr[rr] - k0, e = m.ptr[rr + 1] - k0;
            const double2 x2 = x[rr];
            const double2 d2 = load_stream(&m.diag[rr]);
            double sr = d2.x * x2.x - d2.y * x2.y, si = d2.x * x2.y + d2.y * x2.x;
            for (int k = a; k < e; k++) {
                sr += prod[k].x;
                si += prod[k].y;
            }
            epi.row(rr, make_double2(sr, si), x2, epi.load(rr));
        }
        lds_barrier();
        lo = nlo;
        hi = nhi;
        nlo = nnlo;
        nhi = nnhi;
#pragma unroll
        for (int i = 0; i < kVdNI; i++) {
            c[i] = cn[i];
            vi[i] = vin[i];
        }
    }
}

struct SpmvEpiVd {
    using Pre = int;
    double2 alpha;
    double2* y;
    __device__ __forceinline__ Pre load(int) const { return 0; }
    __device__ __forceinline__ void row(int r, double2 s, double2, Pre) { y[r] = cmul(alpha, s); }
};

__global__ __launch_bounds__(kBlock) void spmv_vd_kernel(VdView m, const double2* __restrict__ x,
                                                         double2* __restrict__ y, double2 alpha) {
    __shared__ double2 lds[kVdLds];
    SpmvEpiVd epi{alpha, y};
    stream_rows_vd(m, x, lds, epi);
}

// As ApplyEpi, with u_j[r] (= x[r]) and u_{j-1}[r] loaded ahead by the row loop.
__global__ __launch_bounds__(kBlock) void lanczos_apply_vd_kernel(LanczosArgs a) {
    KrylovScalars* ks = a.ks;
    if (ks->broken) return;
    __shared__ double2 lds[kVdLds];
    const int j = a.j;
    const double sj = ks->s[j];
    ApplyEpiVd epi;
    epi.c1 = make_double2(a.alpha.x * sj, a.alpha.y * sj);
    epi.c2 = j > 0 ? ks->beta[j - 1] * ks->s[j - 1] : 0.0;
    epi.up = j > 0 ? a.U + (size_t)(j - 1) * a.ld : nullptr;
    epi.w = a.U + (size_t)(j + 1) * a.ld;
    epi.dot = make_double2(0.0, 0.0);
    stream_rows_vd(a.vd, a.U + (size_t)j * a.ld, lds, epi);
    double2 bs = block_sum2<kBlock>(epi.dot, lds);      // the products' LDS is free again
    if (threadIdx.x == 0) a.partials[blockIdx.x] = bs;
    if (last_block_arrive(a.ticket)) {
        double2 tot = reduce_partials<kBlock>(a.partials, gridDim.x, 1, lds);
        if (threadIdx.x == 0) {
            ks->a[j] = make_double2(sj * tot.x, sj * tot.y);
            *a.ticket = 0;
        }
    }
}

// ---------------------------------------------------------------------------------------
// Sliced-ELL value-dictionary kernels: one lane per row, a wavefront per slice of kSlice rows
// (sell_slice_n). The sweep stages the value table in LDS and hands each wavefront its slices
// in the schedule order; the wavefronts of a CU work on different slices, so gathers from x
// overlap without LDS staging or barriers.
// ---------------------------------------------------------------------------------------
template <class Slice>
__device__ __forceinline__ void sell_sweep(const SellView& m, double2* vt, Slice&& slice) {
    for (int i = threadIdx.x; i < m.n_values; i += blockDim.x) vt[i] = m.vtab[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    int s0, end, stride;
    xcd_sweep_waves(m.n_slices, s0, end, stride);
    int kx = 0, kstride = 0;
    const int C = m.chunk;
    if (C > 0 && (gridDim.x & 7) == 0) {               // one front: chunk c belongs to XCD c % 8
        kx = (blockIdx.x >> 3) * (blockDim.x >> 6) + (threadIdx.x >> 6);
        kstride = (gridDim.x >> 3) * (blockDim.x >> 6);
    }
    const int xcd = blockIdx.x & 7;
    kx = __builtin_amdgcn_readfirstlane(kx);
    s0 = __builtin_amdgcn_readfirstlane(s0);          // wave-uniform: slice offsets via scalar loads
    end = __builtin_amdgcn_readfirstlane(end);
    stride = __builtin_amdgcn_readfirstlane(stride);
    for (int k = 0;; k++) {
        int s;
        if (kstride > 0) {
            const int kk = kx + k * kstride;
            s = (8 * (kk / C) + xcd) * C + kk % C;
            if (s >= m.n_slices) break;
        } else {
            s = s0 + k * stride;
            if (s >= end) break;
        }
        s = __builtin_amdgcn_readfirstlane(s);
        slice(s, lane);
    }
}

__global__ __launch_bounds__(kBlock) void spmv_sell_kernel(SellView m, const double2* __restrict__ x,
                                                           double2* __restrict__ y, double2 alpha) {
    __shared__ double2 vt[kVdMaxValues];
    SpmvEpiVd epi{alpha, y};
    sell_sweep(m, vt, [&](int s, int lane) { sell_slice(m, x, vt, s, lane, epi); });
}

__device__ __forceinline__ ApplyEpiVd apply_epi(const LanczosArgs& a) {
    const KrylovScalars* ks = a.ks;
    const int j = a.j;
    const double sj = ks->s[j];
    ApplyEpiVd epi;
    epi.c1 = make_double2(a.alpha.x * sj, a.alpha.y * sj);
    epi.c2 = j > 0 ? ks->beta[j - 1] * ks->s[j - 1] : 0.0;
    epi.up = j > 0 ? a.U + (size_t)(j - 1) * a.ld : nullptr;
    epi.w = a.U + (size_t)(j + 1) * a.ld;
    epi.dot = make_double2(0.0, 0.0);
    return epi;
}

// The finish of an apply: alpha_j = s_j sum of the block partials.
__device__ __forceinline__ void apply_finish(const LanczosArgs& a, double2 tot) {
    const double sj = a.ks->s[a.j];
    a.ks->a[a.j] = make_double2(sj * tot.x, sj * tot.y);
}

// The sliced-ELL apply ends at each block's partial of <u_j, w>; lanczos_finish_kernel sums
// them in block order in a launch of its own. A last-block reduction inside the apply costs
// every block an agent-scope release, which on this multi-XCD part writes back the XCD's L2
// (a one-front stream of the projection's shape, 20 M rows: 240 us with the in-kernel
// reduction, 182 us with partials and a separate one-block finish).
__device__ __forceinline__ void apply_sell_body(const LanczosArgs& a, double2* vt) {
    ApplyEpiVd epi = apply_epi(a);
    const double2* x = a.U + (size_t)a.j * a.ld;
    sell_sweep(a.sell, vt, [&](int s, int lane) { sell_slice(a.sell, x, vt, s, lane, epi); });
    double2 bs = block_sum2<kBlock>(epi.dot, vt);       // the table's LDS is free again
    if (threadIdx.x == 0) a.partials[blockIdx.x] = bs;
}

__global__ __launch_bounds__(kBlock) void lanczos_apply_sell_kernel(LanczosArgs a) {
    if (a.ks->broken) return;
    __shared__ double2 vt[kVdMaxValues];
    apply_sell_body(a, vt);
}

// Two Lanczos recurrences on the same matrix, iteration j of each, in one pass over its words
// (a and b: own bases, scalars and partials). A recurrence that broke down is skipped; with
// one left the single-vector pass runs. Each vector gets the sums and partials of its own
// apply (and, in lanczos_finish_kernel, its reduction order).
__global__ __launch_bounds__(kBlock) void lanczos_apply_sell2_kernel(LanczosArgs a, LanczosArgs b) {
    const bool la = !a.ks->broken, lb = !b.ks->broken;
    __shared__ double2 vt[kVdMaxValues];
    if (!la || !lb) {
        if (la) apply_sell_body(a, vt);
        if (lb) apply_sell_body(b, vt);
        return;
    }
    ApplyEpiVd epi[2] = {apply_epi(a), apply_epi(b)};
    const double2* xs[2] = {a.U + (size_t)a.j * a.ld, b.U + (size_t)b.j * b.ld};
    sell_sweep(a.sell, vt, [&](int s, int lane) { sell_slice_n<2>(a.sell, xs, vt, s, lane, epi); });
    const double2 bs0 = block_sum2<kBlock>(epi[0].dot, vt);
    const double2 bs1 = block_sum2<kBlock>(epi[1].dot, vt);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = bs0;
        b.partials[blockIdx.x] = bs1;
    }
}

// One block, launched after a sliced-ELL apply (PROJECT false: alpha_j) or a projection
// (PROJECT true: beta_j, s_{j+1} or the breakdown) of one recurrence, or of two (`two`: b with
// its own partials): sums the launch's nparts block partials in block order. A recurrence that
// had broken down before that launch wrote no partials and is skipped here too (only this
// kernel sets the flag, after both were read).
template <bool PROJECT>
__global__ __launch_bounds__(kBlock) void lanczos_finish_kernel(LanczosArgs a, LanczosArgs b, int two, int nparts) {
    __shared__ double2 sh[kBlock / 64];
    const bool la = !a.ks->broken, lb = two && !b.ks->broken;
    if (la) {
        const double2 t = reduce_partials<kBlock>(a.partials, nparts, 1, sh);
        if (threadIdx.x == 0) {
            if (PROJECT) project_finish(a, t);
            else apply_finish(a, t);
        }
    }
    if (lb) {
        const double2 t = reduce_partials<kBlock>(b.partials, nparts, 1, sh);
        if (threadIdx.x == 0) {
            if (PROJECT) project_finish(b, t);
            else apply_finish(b, t);
        }
    }
}

// Row blocks on the merge-path coordinate ptr[r] + r (entries plus rows before row r):
// rb[b] = first row with ptr[r] + r >= b * kStreamChunk, rb[nrb] = n. A block then holds at
// most kStreamChunk rows and at most kStreamChunk - 1 + (longest row) entries, so neither
// long rows nor long runs of empty rows unbalance the blocks. Also returns the longest row.
__global__ void row_blocks_kernel(const int* __restrict__ ptr, int64_t n, int nrb, int chunk, int* rb, int* maxlen) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int len = ptr[i + 1] - ptr[i];
        atomicMax(maxlen, len);
    }
    if (i <= nrb) {
        if (i == nrb) {
            rb[i] = (int)n;
        } else {
            const int64_t target = i * (int64_t)chunk;
            int64_t lo = 0, hi = n;                 // first r in [0, n] with ptr[r] + r >= target
            while (lo < hi) {
                const int64_t mid = (lo + hi) / 2;
                if ((int64_t)ptr[mid] + mid >= target) hi = mid;
                else lo = mid + 1;
            }
            rb[i] = (int)lo;
        }
    }
}

__global__ __launch_bounds__(kBlock) void zdotc_kernel(const double2* __restrict__ x, const double2* __restrict__ y,
                                                       int64_t n, double2* partials, unsigned int* ticket,
                                                       double2* result, int slot, int mode) {
    __shared__ double2 sh[kBlock / 64];
    int64_t lo, hi;
    block_rows(n, lo, hi);
    double2 acc = make_double2(0.0, 0.0);
    if (mode == 0) {
        for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
            const double2 u = x[i], v = y[i];
            acc.x += u.x * v.x + u.y * v.y;
            acc.y += u.x * v.y - u.y * v.x;
        }
    } else {
        for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
            const double2 u = x[i];
            acc.x += u.x * u.x + u.y * u.y;
        }
    }
    double2 bs = block_sum2<kBlock>(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
    if (last_block_arrive(ticket)) {
        double2 tot = reduce_partials<kBlock>(partials, gridDim.x, 1, sh);
        if (threadIdx.x == 0) {
            result[slot] = tot;
            *ticket = 0;
        }
    }
}

// ---------------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------------
#define TE_LPR_SWITCH(lpr, MACRO) \
    switch (lpr) {                \
        case 1: MACRO(1); break;  \
        case 2: MACRO(2); break;  \
        case 4: MACRO(4); break;  \
        case 8: MACRO(8); break;  \
        case 16: MACRO(16); break;\
        case 32: MACRO(32); break;\
        default: MACRO(64); break;\
    }

static VdView vd_view(const te_matrix* A) {
    const VdCsr& v = A->vd;
    return VdView{v.ptr, v.col, v.vidx, v.vtab, v.n_values, v.diag, v.blocks, v.n_row_blocks};
}

static SellView sell_view(const te_matrix* A) {
    const VdCsr& v = A->vd;
    return SellView{v.sell_meta, v.sell_word, v.sell_base, v.vtab,      v.n_values, v.diag,
                    v.n_slices,  A->n_rows,   v.sell_chunk, v.diag_re, v.sell_ufix, v.sell_uni};
}

SellView sell_view_of(const te_matrix* A) { return sell_view(A); }

// y = alpha A x in the reference order (a reordered value-dictionary copy is for the evolver).
void launch_spmv(const te_matrix* A, double2 alpha, const double2* x, double2* y, hipStream_t s) {
    if (A->use_vd && !A->vd.perm && A->vd.sell) {
        hipLaunchKernelGGL(spmv_sell_kernel, dim3(A->vd.sell_grid), dim3(kBlock), 0, s, sell_view(A), x, y, alpha);
        return;
    }
    if (A->use_vd && !A->vd.perm) {
        hipLaunchKernelGGL(spmv_vd_kernel, dim3(A->vd.grid), dim3(kBlock), 0, s, vd_view(A), x, y, alpha);
        return;
    }
    const int grid = A->shape.grid;
    if (A->shape.kind == 1) {
        if (A->n_parts > 1) {
            const int P = A->n_parts;
            for (int p = 0; p + 1 < P; p++) {
                const CsrPart& q = A->parts[p];
                hipLaunchKernelGGL(spmv_part_kernel, dim3(grid), dim3(kBlock), 0, s, q.row_block, q.n_row_blocks,
                                   q.ptr, q.col, q.val, x, p > 0 ? A->acc : nullptr, A->acc, nullptr);
            }
            const CsrPart& q = A->parts[P - 1];
            hipLaunchKernelGGL(spmv_stream_kernel, dim3(grid), dim3(kBlock), 0, s, q.row_block, q.n_row_blocks, q.ptr,
                               q.col, q.val, x, y, alpha, A->acc);
            return;
        }
        hipLaunchKernelGGL(spmv_stream_kernel, dim3(grid), dim3(kBlock), 0, s, A->row_block, A->n_row_blocks,
                           A->row_ptr, A->col, A->val, x, y, alpha, nullptr);
        return;
    }
#define L_(LPR) hipLaunchKernelGGL(spmv_kernel<LPR>, dim3(grid), dim3(kBlock), 0, s, (int64_t)A->n_rows, \
                                   A->row_ptr, A->col, A->val, x, y, alpha)
    TE_LPR_SWITCH(A->shape.lanes_per_row, L_)
#undef L_
}

void launch_lanczos_apply(LanczosArgs a, const te_matrix* A, hipStream_t s) {
    if (A->use_vd && A->vd.sell) {
        const VdCsr& v = A->vd;
        a.sell = sell_view(A);
        hipLaunchKernelGGL(lanczos_apply_sell_kernel, dim3(v.sell_grid), dim3(kBlock), 0, s, a);
        hipLaunchKernelGGL(lanczos_finish_kernel<false>, dim3(1), dim3(kBlock), 0, s, a, a, 0, v.sell_grid);
        return;
    }
    if (A->use_vd) {
        a.vd = vd_view(A);
        hipLaunchKernelGGL(lanczos_apply_vd_kernel, dim3(A->vd.grid), dim3(kBlock), 0, s, a);
        return;
    }
    const SpmvShape& sh = A->shape;
    const int grid = sh.grid;
    if (sh.kind == 1) {
        a.acc_in = nullptr;
        if (A->n_parts > 1) {
            const int P = A->n_parts;
            const double2* x = a.U + (size_t)a.j * a.ld;
            for (int p = 0; p + 1 < P; p++) {
                const CsrPart& q = A->parts[p];
                hipLaunchKernelGGL(spmv_part_kernel, dim3(grid), dim3(kBlock), 0, s, q.row_block, q.n_row_blocks,
                                   q.ptr, q.col, q.val, x, p > 0 ? A->acc : nullptr, A->acc, &a.ks->broken);
            }
            const CsrPart& q = A->parts[P - 1];
            a.row_block = q.row_block;
            a.n_row_blocks = q.n_row_blocks;
            a.ptr = q.ptr;
            a.col = q.col;
            a.val = q.val;
            a.acc_in = A->acc;
        }
        hipLaunchKernelGGL(lanczos_apply_stream_kernel, dim3(grid), dim3(kBlock), 0, s, a);
        return;
    }
#define L_(LPR) hipLaunchKernelGGL(lanczos_apply_kernel<LPR>, dim3(grid), dim3(kBlock), 0, s, a)
    TE_LPR_SWITCH(sh.lanes_per_row, L_)
#undef L_
}

// Per row and slice: number of entries with col in [bound[p], bound[p+1]) (columns ascending).
__global__ void split_count_kernel(const int* __restrict__ ptr, const int* __restrict__ col, int64_t n, int P,
                                   int width, int64_t* lens) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int cnt[kMaxParts];
    for (int p = 0; p < P; p++) cnt[p] = 0;
    for (int k = ptr[r]; k < ptr[r + 1]; k++) {
        int p = col[k] / width;
        cnt[p < P ? p : P - 1]++;
    }
    for (int p = 0; p < P; p++) lens[(size_t)p * (n + 1) + r] = cnt[p];
    if (r == 0)
        for (int p = 0; p < P; p++) lens[(size_t)p * (n + 1) + n] = 0;
}

__global__ void split_fill_kernel(const int* __restrict__ ptr, const int* __restrict__ col,
                                  const double2* __restrict__ val, int64_t n, int P, const int64_t* offs,
                                  CsrPart* parts) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > n) return;
    for (int p = 0; p < P; p++) parts[p].ptr[r] = (int)offs[(size_t)p * (n + 1) + r];
    if (r == n) return;
    int k = ptr[r];
    for (int p = 0; p < P; p++) {
        int o = (int)offs[(size_t)p * (n + 1) + r];
        const int e = (int)offs[(size_t)p * (n + 1) + r + 1];
        for (; o < e; o++, k++) {
            parts[p].col[o] = col[k];
            parts[p].val[o] = val[k];
        }
    }
}

static int exclusive_scan64(te_ctx* ctx, const int64_t* in, int64_t* out, int64_t n) {
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

static int row_blocks(te_ctx* ctx, const int* ptr, int64_t n, int64_t nnz, int** rb_out, int* nrb_out, int* maxlen,
                      int chunk = kStreamChunk) {
    const int nrb = (int)((nnz + n + chunk - 1) / chunk);
    int* rb = nullptr;
    TE_HIP(ctx, hipMalloc(&rb, sizeof(int) * (nrb + 1)));
    int* d_max = nullptr;
    TE_HIP(ctx, hipMalloc(&d_max, sizeof(int)));
    TE_HIP(ctx, hipMemsetAsync(d_max, 0, sizeof(int), ctx->stream));
    const int64_t items = std::max<int64_t>(n, nrb + 1);
    hipLaunchKernelGGL(row_blocks_kernel, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, ctx->stream, ptr, n, nrb,
                       chunk, rb, d_max);
    int mx = 0;
    TE_HIP(ctx, hipMemcpyAsync(&mx, d_max, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TE_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(d_max);
    *rb_out = rb;
    *nrb_out = nrb;
    if (maxlen) *maxlen = mx;
    return TE_OK;
}

// Column-split copy (P slices of equal column width) with its own row blocks per slice.
static int build_parts(te_ctx* ctx, te_matrix* A, int P) {
    const int64_t n = A->n_rows;
    const int width = (int)((A->n_cols + P - 1) / P);
    int64_t *lens = nullptr, *offs = nullptr;
    TE_HIP(ctx, hipMalloc(&lens, sizeof(int64_t) * (size_t)P * (n + 1)));
    TE_HIP(ctx, hipMalloc(&offs, sizeof(int64_t) * (size_t)P * (n + 1)));
    hipLaunchKernelGGL(split_count_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, A->row_ptr,
                       A->col, n, P, width, lens);
    for (int p = 0; p < P; p++) {
        int rc = exclusive_scan64(ctx, lens + (size_t)p * (n + 1), offs + (size_t)p * (n + 1), n + 1);
        if (rc != TE_OK) return rc;
    }
    std::vector<int64_t> tot(P);
    for (int p = 0; p < P; p++)
        TE_HIP(ctx, hipMemcpy(&tot[p], offs + (size_t)p * (n + 1) + n, sizeof(int64_t), hipMemcpyDeviceToHost));
    for (int p = 0; p < P; p++) {
        CsrPart& q = A->parts[p];
        q.nnz = tot[p];
        TE_HIP(ctx, hipMalloc(&q.ptr, sizeof(int) * (n + 1)));
        TE_HIP(ctx, hipMalloc(&q.col, sizeof(int) * std::max<int64_t>(q.nnz, 1)));
        TE_HIP(ctx, hipMalloc(&q.val, sizeof(double2) * std::max<int64_t>(q.nnz, 1)));
    }
    CsrPart* d_parts = nullptr;
    TE_HIP(ctx, hipMalloc(&d_parts, sizeof(CsrPart) * P));
    TE_HIP(ctx, hipMemcpy(d_parts, A->parts, sizeof(CsrPart) * P, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(split_fill_kernel, dim3((unsigned)((n + 256) / 256)), dim3(256), 0, ctx->stream, A->row_ptr,
                       A->col, A->val, n, P, offs, d_parts);
    TE_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(d_parts);
    hipFree(lens);
    hipFree(offs);
    for (int p = 0; p < P; p++) {
        CsrPart& q = A->parts[p];
        int rc = row_blocks(ctx, q.ptr, n, q.nnz, &q.row_block, &q.n_row_blocks, nullptr);
        if (rc != TE_OK) return rc;
    }
    TE_HIP(ctx, hipMalloc(&A->acc, sizeof(double2) * n));
    A->n_parts = P;
    return TE_OK;
}

void free_parts(te_matrix* A) {
    for (int p = 0; p < kMaxParts; p++) {
        CsrPart& q = A->parts[p];
        if (q.ptr) hipFree(q.ptr);
        if (q.col) hipFree(q.col);
        if (q.val) hipFree(q.val);
        if (q.row_block) hipFree(q.row_block);
        q = CsrPart{};
    }
    if (A->acc) hipFree(A->acc);
    A->acc = nullptr;
    A->n_parts = 0;
}

// ---------------------------------------------------------------------------------------
// Value-dictionary build: distinct off-diagonal values found with a lock-free open-addressing
// set of 64-bit keys (every probe loop is bounded by the slot count), then checked bit for
// bit against the table; any overflow, key collision or repeated diagonal entry keeps the
// plain CSR kernels.
// ---------------------------------------------------------------------------------------
constexpr int kVdSlots = 4096;      // power of two, >= 4 kVdMaxValues

__device__ __forceinline__ unsigned long long vd_bits(double v) { return (unsigned long long)__double_as_longlong(v); }

__device__ __forceinline__ unsigned long long vd_key(double2 v) {
    unsigned long long h = vd_bits(v.x) * 0x9E3779B97F4A7C15ull;
    h ^= vd_bits(v.y) + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2);
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    return h | 1ull;                // 0 marks an empty slot
}

enum { kVdFlagRepeatedDiag = 1, kVdFlagFull = 2, kVdFlagMismatch = 4 };

// Per row of the copy (old row perm[r], or r): off-diagonal count; flags[2] = longest such row.
__global__ void vd_count_kernel(const int* __restrict__ ptr, const int* __restrict__ col, const int* __restrict__ perm,
                                int64_t n, int64_t* offlen, int* flags) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > n) return;
    if (r == n) {
        offlen[n] = 0;
        return;
    }
    const int o = perm ? perm[r] : (int)r;
    int nd = 0;
    for (int k = ptr[o]; k < ptr[o + 1]; k++) nd += col[k] == o;
    const int len = ptr[o + 1] - ptr[o] - nd;
    offlen[r] = len;
    atomicMax(&flags[2], len);
    if (nd > 1) atomicOr(flags, (int)kVdFlagRepeatedDiag);
}

__global__ void vd_insert_kernel(const int* __restrict__ ptr, const int* __restrict__ col,
                                 const double2* __restrict__ val, int64_t n, unsigned long long* keys,
                                 double2* slot_val, int* flags) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int k = ptr[r]; k < ptr[r + 1]; k++) {
        if (col[k] == (int)r) continue;
        const double2 v = val[k];
        const unsigned long long key = vd_key(v);
        int s = (int)(key & (kVdSlots - 1));
        bool placed = false;
        for (int probe = 0; probe < kVdSlots; probe++) {
            unsigned long long cur = __hip_atomic_load(&keys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == 0ull) {
                cur = atomicCAS(&keys[s], 0ull, key);
                if (cur == 0ull) slot_val[s] = v;     // this thread created the slot
            }
            if (cur == 0ull || cur == key) {
                placed = true;
                break;
            }
            s = (s + 1) & (kVdSlots - 1);
        }
        if (!placed) atomicOr(flags, (int)kVdFlagFull);
    }
}

// Dense table in slot order (one thread: 4096 slots).
__global__ void vd_compact_kernel(const unsigned long long* keys, const double2* slot_val, int* dense, double2* vtab,
                                  int* n_values) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int c = 0;
    for (int s = 0; s < kVdSlots; s++) {
        if (keys[s] != 0ull) {
            dense[s] = c;
            if (c < kVdMaxValues) vtab[c] = slot_val[s];
            c++;
        } else {
            dense[s] = -1;
        }
    }
    *n_values = c;
}

// Row r of the copy = old row perm[r] (or r), columns mapped through iperm and sorted.
__global__ void vd_fill_kernel(const int* __restrict__ ptr, const int* __restrict__ col,
                               const double2* __restrict__ val, const int* __restrict__ perm,
                               const int* __restrict__ iperm, int64_t n, const int64_t* __restrict__ offs,
                               const unsigned long long* __restrict__ keys, const double2* __restrict__ slot_val,
                               const int* __restrict__ dense, int* vptr, int* vcol, unsigned short* vidx,
                               double2* diag, int* flags) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > n) return;
    vptr[r] = (int)offs[r];
    if (r == n) return;
    const int o0 = (int)offs[r];
    int o = o0;
    const int orow = perm ? perm[r] : (int)r;
    double2 d = make_double2(0.0, 0.0);
    for (int k = ptr[orow]; k < ptr[orow + 1]; k++) {
        const int c = col[k];
        const double2 v = val[k];
        if (c == orow) {
            d = v;
            continue;
        }
        const unsigned long long key = vd_key(v);
        int s = (int)(key & (kVdSlots - 1)), found = -1;
        for (int probe = 0; probe < kVdSlots; probe++) {
            const unsigned long long cur = keys[s];
            if (cur == key) {
                found = s;
                break;
            }
            if (cur == 0ull) break;
            s = (s + 1) & (kVdSlots - 1);
        }
        int idx = 0;
        if (found < 0 || vd_bits(slot_val[found].x) != vd_bits(v.x) || vd_bits(slot_val[found].y) != vd_bits(v.y))
            atomicOr(flags, (int)kVdFlagMismatch);
        else
            idx = dense[found];
        vcol[o] = iperm ? iperm[c] : c;
        vidx[o] = (unsigned short)idx;
        o++;
    }
    diag[r] = d;
    if (iperm) {                         // insertion sort of the row by new column (rows are short)
        for (int a = o0 + 1; a < o; a++) {
            const int cc = vcol[a];
            const unsigned short vv = vidx[a];
            int b = a - 1;
            while (b >= o0 && vcol[b] > cc) {
                vcol[b + 1] = vcol[b];
                vidx[b + 1] = vidx[b];
                b--;
            }
            vcol[b + 1] = cc;
            vidx[b + 1] = vv;
        }
    }
}

__global__ void vd_blocks_kernel(const int* __restrict__ rb, const int* __restrict__ ptr, int nrb, int2* blocks) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= nrb) blocks[i] = make_int2(rb[i], ptr[rb[i]]);
}

__global__ void invert_perm_kernel(const int* __restrict__ perm, int64_t n, int* iperm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) iperm[perm[i]] = (int)i;
}

__global__ void gather_kernel(const double2* __restrict__ src, const int* __restrict__ idx, double2* __restrict__ dst,
                              int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

__global__ void gather_re_kernel(const double* __restrict__ src, const int* __restrict__ idx, double* __restrict__ dst,
                                 int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

static unsigned gather_grid(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 8192); }

void launch_gather(const double2* src, const int* idx, double2* dst, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(gather_kernel, dim3(gather_grid(n)), dim3(256), 0, s, src, idx, dst, n);
}

void launch_gather_re(const double* src, const int* idx, double* dst, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(gather_re_kernel, dim3(gather_grid(n)), dim3(256), 0, s, src, idx, dst, n);
}

// Within each window of kSellWindow rows of the base order, rows by decreasing off-diagonal
// count (ties in base order): out[w0 + rank] = base[w0 + i] (or w0 + i).
__global__ __launch_bounds__(kSellWindow) void sigma_sort_kernel(const int64_t* __restrict__ len, int64_t n,
                                                                 const int* __restrict__ base, int* out) {
    __shared__ int sl[kSellWindow];
    const int64_t w0 = (int64_t)blockIdx.x * kSellWindow;
    const int i = threadIdx.x;
    const int64_t r = w0 + i;
    const int cnt = (int)(n - w0 < kSellWindow ? n - w0 : kSellWindow);
    sl[i] = r < n ? (int)len[r] : -1;
    __syncthreads();
    if (r >= n) return;
    const int li = sl[i];
    int rank = 0;
    for (int j2 = 0; j2 < cnt; j2++) {
        const int lj = sl[j2];
        rank += (lj > li) || (lj == li && j2 < i);
    }
    out[w0 + rank] = base ? base[r] : (int)r;
}

// What the two passes of sell_align_kernel read and write.
struct SellBuild {
    const int* vptr;
    const int* vcol;
    const unsigned short* vvidx;
    int64_t n;
    int n_slices;
    int shared;                 // 1: look for shared slots
    // count pass
    int64_t* slen;              // middle slots * kSlice per slice
    int64_t* smax;              // the slice's longest middle part * kSlice
    int64_t* ucnt;              // shared records per slice
    int* usplit;                // leading shared records per slice
    int64_t* ovcnt;             // overflow records per slice
    // fill pass
    const int64_t* soff;        // exclusive scan of slen
    const int64_t* ovoff;       // exclusive scan of ovcnt
    int4* meta;
    unsigned* word;
    int* base;
    int2* ufix;                 // null: no shared slots anywhere
    int2* uni;
};

// Shared slots of a full slice (all kSlice rows exist): the leading entry positions 0..ulo-1 and
// the trailing positions deg-uhi..deg-1 at which every row of the slice has an entry with the same
// column offset from its row and the same value index (a transition that changes only modes the
// slice's rows share). They are stored once per slice; the entries between them (the middle) go
// to per-row slots. Each row's entries keep their column order: leading, middle, trailing.
__device__ __forceinline__ void shared_slots(const SellBuild& b, int64_t r, int k, int kend, int& ulo, int& uhi) {
    ulo = uhi = 0;
    int mn = kend - k;
#pragma unroll
    for (int d = kSlice / 2; d >= 1; d >>= 1) mn = min(mn, __shfl_xor(mn, d));
    auto same = [&](int e) {
        const int off = b.vcol[e] - (int)r, vi = b.vvidx[e];
        return __all(off == __shfl(off, 0) && vi == __shfl(vi, 0)) != 0;
    };
    while (ulo < mn && same(k + ulo)) ulo++;
    while (ulo + uhi < mn && same(kend - 1 - uhi)) uhi++;
}

// Slot alignment of the slices, one wavefront per slice (lane l: row s*kSlice + l). After the
// shared slots, the next slot of a slice takes, from every row, the row's next middle entry (rows
// in column order) if that lies within kSellSpan columns of the smallest next entry over the
// slice; rows whose next entry lies further right get a padding word in it. Count pass (FILL
// false): middle slots, longest middle part (the slots without the span limit) and shared records
// per slice. Fill pass: the words and slot bases at the offsets soff, the shared records at uoff.
// A slot's padding words point at its base column, a column of one of its entries.
template <bool FILL>
__global__ __launch_bounds__(256) void sell_align_kernel(SellBuild b) {
    const int s = blockIdx.x * (blockDim.x / kSlice) + (int)(threadIdx.x / kSlice);
    const int lane = threadIdx.x % kSlice;
    if (s > b.n_slices) return;                                // whole wavefronts
    if (s == b.n_slices) {
        if (!FILL && lane == 0) b.slen[s] = b.smax[s] = b.ucnt[s] = b.ovcnt[s] = 0;
        return;
    }
    const int64_t r = (int64_t)s * kSlice + lane;
    int k = r < b.n ? b.vptr[r] : 0;
    int kend = r < b.n ? b.vptr[r + 1] : 0;
    int ulo = 0, uhi = 0;
    if (FILL) {
        ulo = b.usplit[s];
        uhi = (int)b.ucnt[s] - ulo;
    } else if (b.shared && (int64_t)(s + 1) * kSlice <= b.n) {
        shared_slots(b, r, k, kend, ulo, uhi);
    }
    if (FILL && lane == 0) {
        const int nsh = ulo + uhi;
        const int64_t ov = b.ovoff[s];
        b.meta[s] = make_int4((int)b.soff[s], (int)ov, ulo | (nsh << 16), 0);
        if (s == b.n_slices - 1) b.meta[b.n_slices] = make_int4((int)b.soff[b.n_slices], (int)b.ovoff[b.n_slices], 0, 0);
        if (b.ufix) {
            for (int q = 0; q < kSellFix || q < nsh; q++) {
                int2 rec = make_int2(0, 0);
                if (q < nsh) {
                    const int e = q < ulo ? k + q : kend - uhi + (q - ulo);
                    rec = make_int2(b.vcol[e] - (int)r, (int)b.vvidx[e]);
                }
                if (q < kSellFix) b.ufix[(int64_t)s * kSellFix + q] = rec;
                else b.uni[ov + q - kSellFix] = rec;
            }
        }
    }
    k += ulo;
    kend -= uhi;
    if (!FILL) {
        int len = kend - k;
#pragma unroll
        for (int d = kSlice / 2; d >= 1; d >>= 1) len = max(len, __shfl_xor(len, d));
        if (lane == 0) {
            b.smax[s] = (int64_t)len * kSlice;
            b.ucnt[s] = ulo + uhi;
            b.usplit[s] = ulo;
            b.ovcnt[s] = ulo + uhi > kSellFix ? ulo + uhi - kSellFix : 0;
        }
    }
    const int64_t o = FILL ? b.soff[s] : 0;
    for (int j = 0;; j++) {
        const bool has = k < kend;
        const int c = has ? b.vcol[k] : INT_MAX;
        int lo = c;
#pragma unroll
        for (int d = kSlice / 2; d >= 1; d >>= 1) lo = min(lo, __shfl_xor(lo, d));
        if (lo == INT_MAX) {                                   // the same for all lanes
            if (!FILL && lane == 0) b.slen[s] = (int64_t)j * kSlice;
            return;
        }
        const bool take = has && (unsigned)(c - lo) < (unsigned)kSellSpan;
        if (FILL) {
            b.word[o + (int64_t)j * kSlice + lane] =
                take ? ((unsigned)b.vvidx[k] << 16) | (unsigned)(c - lo) : (kSellPadIdx << 16);
            if (lane == 0) b.base[o / kSlice + j] = lo;
        }
        k += take ? 1 : 0;
    }
}

void free_vd(te_matrix* A) {
    VdCsr& v = A->vd;
    if (v.ptr) hipFree(v.ptr);
    if (v.col) hipFree(v.col);
    if (v.vidx) hipFree(v.vidx);
    if (v.vtab) hipFree(v.vtab);
    if (v.diag) hipFree(v.diag);
    if (v.diag_re) hipFree(v.diag_re);
    if (v.blocks) hipFree(v.blocks);
    if (v.perm) hipFree(v.perm);
    if (v.iperm) hipFree(v.iperm);
    if (v.sell_meta) hipFree(v.sell_meta);
    if (v.sell_word) hipFree(v.sell_word);
    if (v.sell_base) hipFree(v.sell_base);
    if (v.sell_ufix) hipFree(v.sell_ufix);
    if (v.sell_uni) hipFree(v.sell_uni);
    v = VdCsr{};
    A->use_vd = false;
}

// Sliced-ELL form of A->vd (rows of at most 255 off-diagonal entries).
// re[i] = d[i].x; *complex_flag = 1 if any d[i].y is not +0.0 (bit pattern 0).
__global__ void diag_real_kernel(const double2* __restrict__ d, int64_t n, double* __restrict__ re, int* complex_flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 v = d[i];
    re[i] = v.x;
    if (__double_as_longlong(v.y) != 0) *complex_flag = 1;
}

// The real diagonal of a sliced-ELL matrix (VdCsr::diag_re), when it has one.
static int build_diag_re(te_ctx* ctx, VdCsr& v, int64_t n) {
    if (v.diag_re) hipFree(v.diag_re);
    v.diag_re = nullptr;
    if (n <= 0) return TE_OK;
    double* re = nullptr;
    int* flag = nullptr;
    TE_HIP(ctx, hipMalloc(&re, sizeof(double) * n));
    if (hipMalloc(&flag, sizeof(int)) != hipSuccess) {
        hipFree(re);
        return fail(ctx, TE_ERR_HIP, "hipMalloc failed for the diagonal flag");
    }
    int h = 0;
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(diag_real_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, v.diag, n, re,
                           flag);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(flag);
    if (e != hipSuccess || h != 0) {
        hipFree(re);
        if (e != hipSuccess) return fail(ctx, TE_ERR_HIP, "real-diagonal check failed: %s", hipGetErrorString(e));
        return TE_OK;
    }
    v.diag_re = re;
    return TE_OK;
}

static int build_sell(te_ctx* ctx, te_matrix* A) {
    VdCsr& v = A->vd;
    hipStream_t st = ctx->stream;
    const int64_t n = A->n_rows;
    const int64_t ns64 = (n + kSlice - 1) / kSlice;
    if (ns64 >= INT32_MAX / kSlice) return TE_OK;
    const int ns = (int)ns64;
    int64_t *slen = nullptr, *soff = nullptr, *smax = nullptr, *smax_off = nullptr, *ucnt = nullptr, *uoff = nullptr;
    int64_t *ovcnt = nullptr, *ovoff = nullptr;
    int* usplit = nullptr;
    TE_HIP(ctx, hipMalloc(&ovcnt, sizeof(int64_t) * (ns + 1)));
    TE_HIP(ctx, hipMalloc(&ovoff, sizeof(int64_t) * (ns + 1)));
    TE_HIP(ctx, hipMalloc(&slen, sizeof(int64_t) * (ns + 1)));
    TE_HIP(ctx, hipMalloc(&soff, sizeof(int64_t) * (ns + 1)));
    TE_HIP(ctx, hipMalloc(&smax, sizeof(int64_t) * (ns + 1)));
    TE_HIP(ctx, hipMalloc(&smax_off, sizeof(int64_t) * (ns + 1)));
    TE_HIP(ctx, hipMalloc(&ucnt, sizeof(int64_t) * (ns + 1)));
    TE_HIP(ctx, hipMalloc(&uoff, sizeof(int64_t) * (ns + 1)));
    TE_HIP(ctx, hipMalloc(&usplit, sizeof(int) * (ns + 1)));
    const unsigned ga = (unsigned)((ns + 1 + 256 / kSlice - 1) / (256 / kSlice));
    SellBuild sb{};
    sb.vptr = v.ptr;
    sb.vcol = v.col;
    sb.vvidx = v.vidx;
    sb.n = n;
    sb.n_slices = ns;
    const char* shared_env = std::getenv("TE_SELL_SHARED");          // 0: every slot per row
    sb.shared = !(shared_env && shared_env[0] == '0');
    sb.slen = slen;
    sb.smax = smax;
    sb.ucnt = ucnt;
    sb.usplit = usplit;
    sb.ovcnt = ovcnt;
    hipLaunchKernelGGL(sell_align_kernel<false>, dim3(ga), dim3(256), 0, st, sb);
    int rc = exclusive_scan64(ctx, slen, soff, ns + 1);
    if (rc == TE_OK) rc = exclusive_scan64(ctx, smax, smax_off, ns + 1);
    if (rc == TE_OK) rc = exclusive_scan64(ctx, ucnt, uoff, ns + 1);
    if (rc == TE_OK) rc = exclusive_scan64(ctx, ovcnt, ovoff, ns + 1);
    int64_t total = 0, total_max = 0, n_uni = 0, n_ov = 0;
    if (rc == TE_OK) {
        TE_HIP(ctx, hipMemcpyAsync(&total, soff + ns, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        TE_HIP(ctx, hipMemcpyAsync(&total_max, smax_off + ns, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        TE_HIP(ctx, hipMemcpyAsync(&n_uni, uoff + ns, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        TE_HIP(ctx, hipMemcpyAsync(&n_ov, ovoff + ns, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        TE_HIP(ctx, hipStreamSynchronize(st));
    }
    hipFree(smax);
    hipFree(smax_off);
    hipFree(uoff);
    hipFree(ovcnt);
    // Rows whose columns scatter beyond the span (no locality) would make the aligned slots pad
    // without bound: past 1.5x the slots of a layout padded only to each slice's longest row, the
    // matrix keeps the value-dictionary row-block kernels.
    if (rc == TE_OK && total < INT32_MAX && n_ov < INT32_MAX && total <= total_max + total_max / 2) {
        TE_HIP(ctx, hipMalloc(&v.sell_meta, sizeof(int4) * (ns + 1)));
        if (n_uni > 0) {
            TE_HIP(ctx, hipMalloc(&v.sell_ufix, sizeof(int2) * (size_t)ns * kSellFix));
            TE_HIP(ctx, hipMalloc(&v.sell_uni, sizeof(int2) * std::max<int64_t>(n_ov, 1)));
        }
        // a slice's last group of slots may run up to kSellGroup - 1 slots past its end: zero
        // padding (words: column base + 0, value index 0, never added)
        const int64_t pad = (int64_t)kSellGroup * kSlice;
        const int64_t nslots = total / kSlice;
        TE_HIP(ctx, hipMalloc(&v.sell_word, sizeof(unsigned) * (total + pad)));
        TE_HIP(ctx, hipMalloc(&v.sell_base, sizeof(int) * (nslots + kSellGroup)));
        TE_HIP(ctx, hipMemsetAsync(v.sell_word + total, 0, sizeof(unsigned) * pad, st));
        TE_HIP(ctx, hipMemsetAsync(v.sell_base + nslots, 0, sizeof(int) * kSellGroup, st));
        sb.soff = soff;
        sb.ovoff = ovoff;
        sb.meta = v.sell_meta;
        sb.word = v.sell_word;
        sb.base = v.sell_base;
        sb.ufix = v.sell_ufix;                             // null without shared slots
        sb.uni = v.sell_uni;
        hipLaunchKernelGGL(sell_align_kernel<true>, dim3(ga), dim3(256), 0, st, sb);
        TE_HIP(ctx, hipStreamSynchronize(st));
        v.n_slices = ns;
        v.sell_entries = total;
        v.sell_shared = n_uni;
        v.sell_overflow = n_ov;
        // Three resident blocks per CU (12 waves): the waves of an XCD sweep its slices together,
        // and a narrow window of slices in flight keeps more of x's lines in the XCD's L2. On the
        // L = N = 14 ring (apply, µs): 1296, 826, 806, 867, 885 for 1..5 blocks per CU (5 is the
        // occupancy limit); groups of 8 slots in place of 4 were slower at every count.
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lanczos_apply_sell_kernel, kBlock, 0) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        const char* gmul = std::getenv("TE_SELL_BLOCKS_PER_CU");    // launch-shape sweeps
        const int want = gmul ? std::max(1, std::atoi(gmul)) : 3;
        const int cap = std::max(8, ctx->num_cus * std::min(want, per_cu));
        const int blocks = (ns + kBlock / 64 - 1) / (kBlock / 64);
        v.sell_grid = (int)((std::min(blocks, cap) + 7) / 8 * 8);
        // All XCDs sweep one front, 512 slices (32,768 rows) per chunk: on the L = N = 14 ring
        // 6 % faster than a contiguous eighth per XCD (the far hops' windows of x are then shared
        // in the MALL instead of eight separate ones); apply 383.0-383.5 ms a run at 512 slices
        // a chunk, 385.0-385.4 at 256, 384.1 at 1024, 385.2 at 128.
        const char* chunk = std::getenv("TE_SELL_CHUNK");          // sweep-schedule sweeps
        v.sell_chunk = chunk ? std::max(0, std::atoi(chunk)) : 512;
        v.sell = true;
        rc = build_diag_re(ctx, v, n);
    }
    hipFree(slen);
    hipFree(soff);
    hipFree(ucnt);
    hipFree(ovoff);
    hipFree(usplit);
    return rc;
}

// Builds A->vd from the CSR on the device. Returns TE_OK also when the matrix does not qualify
// (A->use_vd stays false); errors only for HIP failures.
static int build_vd(te_ctx* ctx, te_matrix* A, const int* perm) {
    free_vd(A);
    const int64_t n = A->n_rows;
    if (n != A->n_cols || n == 0) return TE_OK;
    hipStream_t st = ctx->stream;
    int64_t *offlen = nullptr, *offs = nullptr;
    unsigned long long* keys = nullptr;
    double2* slot_val = nullptr;
    int *dense = nullptr, *flags = nullptr;     // flags[0]: kVdFlag*, flags[1]: n_values, flags[2]: longest row
    TE_HIP(ctx, hipMalloc(&offlen, sizeof(int64_t) * (n + 1)));
    TE_HIP(ctx, hipMalloc(&offs, sizeof(int64_t) * (n + 1)));
    TE_HIP(ctx, hipMalloc(&keys, sizeof(unsigned long long) * kVdSlots));
    TE_HIP(ctx, hipMalloc(&slot_val, sizeof(double2) * kVdSlots));
    TE_HIP(ctx, hipMalloc(&dense, sizeof(int) * kVdSlots));
    TE_HIP(ctx, hipMalloc(&flags, sizeof(int) * 3));
    VdCsr& v = A->vd;
    TE_HIP(ctx, hipMalloc(&v.vtab, sizeof(double2) * kVdMaxValues));
    const char* sell_env = std::getenv("TE_SPMV_SELL");              // 0: row-block kernel only
    const bool want_sell = !(sell_env && sell_env[0] == '0');
    const char* sigma_env = std::getenv("TE_SELL_SIGMA");            // 1: length sort in windows
    const bool want_sigma = want_sell && sigma_env && sigma_env[0] == '1';
    const unsigned g1 = (unsigned)((n + 256) / 256);
    if (perm || want_sigma) {
        TE_HIP(ctx, hipMalloc(&v.perm, sizeof(int) * n));
        TE_HIP(ctx, hipMalloc(&v.iperm, sizeof(int) * n));
        if (want_sigma) {       // refine the base order: rows by decreasing length within windows
            TE_HIP(ctx, hipMemsetAsync(flags, 0, sizeof(int) * 3, st));
            hipLaunchKernelGGL(vd_count_kernel, dim3(g1), dim3(256), 0, st, A->row_ptr, A->col, perm, n, offlen, flags);
            hipLaunchKernelGGL(sigma_sort_kernel, dim3((unsigned)((n + kSellWindow - 1) / kSellWindow)),
                               dim3(kSellWindow), 0, st, offlen, n, perm, v.perm);
        } else {
            TE_HIP(ctx, hipMemcpyAsync(v.perm, perm, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
        }
        hipLaunchKernelGGL(invert_perm_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, v.perm, n, v.iperm);
    }
    TE_HIP(ctx, hipMemsetAsync(keys, 0, sizeof(unsigned long long) * kVdSlots, st));
    TE_HIP(ctx, hipMemsetAsync(flags, 0, sizeof(int) * 3, st));
    hipLaunchKernelGGL(vd_count_kernel, dim3(g1), dim3(256), 0, st, A->row_ptr, A->col, v.perm, n, offlen, flags);
    hipLaunchKernelGGL(vd_insert_kernel, dim3(g1), dim3(256), 0, st, A->row_ptr, A->col, A->val, n, keys, slot_val,
                       flags);
    hipLaunchKernelGGL(vd_compact_kernel, dim3(1), dim3(64), 0, st, keys, slot_val, dense, v.vtab, flags + 1);
    int rc = exclusive_scan64(ctx, offlen, offs, n + 1);
    int h[3] = {0, 0, 0};
    int64_t nnz_off = 0;
    if (rc == TE_OK) {
        TE_HIP(ctx, hipMemcpyAsync(h, flags, sizeof(int) * 3, hipMemcpyDeviceToHost, st));
        TE_HIP(ctx, hipMemcpyAsync(&nnz_off, offs + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        TE_HIP(ctx, hipStreamSynchronize(st));
    }
    // rows longer than a row block can hold are not stored (nor sorted, which is quadratic)
    bool ok = rc == TE_OK && h[0] == 0 && h[1] <= kVdMaxValues && h[2] + kVdChunk - 1 <= kVdCap;
    if (ok) {
        v.n_values = h[1];
        v.nnz = nnz_off;
        TE_HIP(ctx, hipMalloc(&v.ptr, sizeof(int) * (n + 1)));
        TE_HIP(ctx, hipMalloc(&v.col, sizeof(int) * (nnz_off + kVdPad)));
        TE_HIP(ctx, hipMalloc(&v.vidx, sizeof(unsigned short) * (nnz_off + kVdPad)));
        TE_HIP(ctx, hipMemsetAsync(v.col + nnz_off, 0, sizeof(int) * kVdPad, st));
        TE_HIP(ctx, hipMemsetAsync(v.vidx + nnz_off, 0, sizeof(unsigned short) * kVdPad, st));
        TE_HIP(ctx, hipMalloc(&v.diag, sizeof(double2) * n));
        hipLaunchKernelGGL(vd_fill_kernel, dim3(g1), dim3(256), 0, st, A->row_ptr, A->col, A->val, v.perm, v.iperm, n,
                           offs, keys, slot_val, dense, v.ptr, v.col, v.vidx, v.diag, flags);
        TE_HIP(ctx, hipMemcpyAsync(h, flags, sizeof(int), hipMemcpyDeviceToHost, st));
        TE_HIP(ctx, hipStreamSynchronize(st));
        ok = h[0] == 0;
    }
    hipFree(offlen);
    hipFree(offs);
    hipFree(keys);
    hipFree(slot_val);
    hipFree(dense);
    hipFree(flags);
    if (rc != TE_OK) return rc;
    int maxlen = 0;
    if (ok) {
        int* rb = nullptr;
        rc = row_blocks(ctx, v.ptr, n, v.nnz, &rb, &v.n_row_blocks, &maxlen, kVdChunk);
        if (rc != TE_OK) return rc;
        ok = v.n_row_blocks > 0 && maxlen + kVdChunk - 1 <= kVdCap;
        if (ok) {
            TE_HIP(ctx, hipMalloc(&v.blocks, sizeof(int2) * (v.n_row_blocks + 1)));
            hipLaunchKernelGGL(vd_blocks_kernel, dim3((unsigned)((v.n_row_blocks + 256) / 256)), dim3(256), 0, st, rb,
                               v.ptr, v.n_row_blocks, v.blocks);
            TE_HIP(ctx, hipStreamSynchronize(st));
        }
        hipFree(rb);
    }
    if (!ok) {
        free_vd(A);
        return TE_OK;
    }
    const char* gmul = std::getenv("TE_VD_BLOCKS_PER_CU");          // launch-shape sweeps
    const int cap = std::max(8, ctx->num_cus * (gmul ? std::max(1, std::atoi(gmul)) : 4));
    v.grid = (int)((std::min(v.n_row_blocks, cap) + 7) / 8 * 8);
    A->use_vd = true;
    if (want_sell && maxlen <= 255) {
        rc = build_sell(ctx, A);
        if (rc != TE_OK) return rc;
    }
    return TE_OK;
}

int prepare_spmv(te_ctx* ctx, te_matrix* A, const int* perm) {
    A->shape = choose_shape(ctx, A->n_rows, A->nnz);
    const int nrb = (int)((A->nnz + A->n_rows + kStreamChunk - 1) / kStreamChunk);
    if (A->row_block) hipFree(A->row_block);
    A->row_block = nullptr;
    TE_HIP(ctx, hipMalloc(&A->row_block, sizeof(int) * (nrb + 1)));
    int* d_max = nullptr;
    TE_HIP(ctx, hipMalloc(&d_max, sizeof(int)));
    TE_HIP(ctx, hipMemsetAsync(d_max, 0, sizeof(int), ctx->stream));
    const int64_t items = std::max<int64_t>(A->n_rows, nrb + 1);
    hipLaunchKernelGGL(row_blocks_kernel, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, ctx->stream, A->row_ptr,
                       (int64_t)A->n_rows, nrb, kStreamChunk, A->row_block, d_max);
    int mx = 0;
    TE_HIP(ctx, hipMemcpyAsync(&mx, d_max, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TE_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(d_max);
    A->n_row_blocks = nrb;
    A->max_row_len = mx;
    const char* force = std::getenv("TE_SPMV_KIND");
    const bool fits = mx + kStreamChunk - 1 <= kStreamCap && nrb > 0;
    if (fits && !(force && force[0] == '0')) {
        A->shape.kind = 1;
        const char* gmul = std::getenv("TE_SPMV_BLOCKS_PER_CU");   // launch-shape sweeps
        const int cap = std::max(8, ctx->num_cus * (gmul ? std::max(1, std::atoi(gmul)) : 8));
        A->shape.grid = (int)((std::min(nrb, cap) + 7) / 8 * 8);
        // Diagonal matrices (observables of n_k type) are read as dense diagonals by the evolver.
        const char* vd = std::getenv("TE_SPMV_VD");                    // 0: plain CSR stream kernels
        if (!(vd && vd[0] == '0') && !A->diagonal) {
            int rc = build_vd(ctx, A, perm);
            if (rc != TE_OK) return rc;
            if (A->use_vd) return TE_OK;
        }
        // Column slices of x no larger than TE_SPMV_PART_MB (MiB). Off by default: measured on
        // the L = N = 14 ring (x = 321 MB), 2, 3 and 5 slices were 10-30 % slower than one
        // pass (the slices are not retained by the MALL under this access pattern).
        const char* mb = std::getenv("TE_SPMV_PART_MB");
        const double budget = (mb ? std::atof(mb) : 0.0) * 1048576.0;
        const double xbytes = 16.0 * (double)A->n_cols;
        int P = budget > 0 ? (int)std::ceil(xbytes / budget) : 1;
        P = std::min(P, kMaxParts);
        if (P > 1) {
            free_parts(A);
            int rc = build_parts(ctx, A, P);
            if (rc != TE_OK) return rc;
            for (int p = 0; p < P; p++)              // every pass must visit every row
                if (A->parts[p].n_row_blocks == 0) {
                    free_parts(A);
                    break;
                }
        }
    }
    return TE_OK;
}

void launch_lanczos_project(const LanczosArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(lanczos_project_kernel, dim3(grid), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(lanczos_finish_kernel<true>, dim3(1), dim3(kBlock), 0, s, a, a, 0, grid);
}

void launch_lanczos_apply2(LanczosArgs a, LanczosArgs b, const te_matrix* A, int grid, hipStream_t s) {
    if (A->use_vd && A->vd.sell) {
        a.sell = b.sell = sell_view(A);
        const int g = grid > 0 ? grid : A->vd.sell_grid;
        hipLaunchKernelGGL(lanczos_apply_sell2_kernel, dim3(g), dim3(kBlock), 0, s, a, b);
        hipLaunchKernelGGL(lanczos_finish_kernel<false>, dim3(1), dim3(kBlock), 0, s, a, b, 1, g);
        return;
    }
    launch_lanczos_apply(a, A, s);
    launch_lanczos_apply(b, A, s);
}

void launch_lanczos_project2(const LanczosArgs& a, const LanczosArgs& b, int grid, hipStream_t s) {
    hipLaunchKernelGGL(lanczos_project2_kernel, dim3(grid), dim3(kBlock), 0, s, a, b);
    hipLaunchKernelGGL(lanczos_finish_kernel<true>, dim3(1), dim3(kBlock), 0, s, a, b, 1, grid);
}

void launch_zdotc(const double2* x, const double2* y, int64_t n, ReduceScratch& rs, int slot, hipStream_t s) {
    hipLaunchKernelGGL(zdotc_kernel, dim3(rs.grid), dim3(kBlock), 0, s, x, y, n, rs.partials, rs.ticket, rs.result,
                       slot, 0);
}

void launch_nrm2sq(const double2* x, int64_t n, ReduceScratch& rs, int slot, hipStream_t s) {
    hipLaunchKernelGGL(zdotc_kernel, dim3(rs.grid), dim3(kBlock), 0, s, x, x, n, rs.partials, rs.ticket, rs.result,
                       slot, 1);
}

int reduce_scratch_alloc(te_ctx* ctx, ReduceScratch& rs, int grid) {
    rs.grid = grid;
    TE_HIP(ctx, hipMalloc(&rs.partials, sizeof(double2) * (size_t)grid * 128));
    TE_HIP(ctx, hipMalloc(&rs.ticket, sizeof(unsigned int) * 4));
    TE_HIP(ctx, hipMalloc(&rs.result, sizeof(double2) * 64));
    TE_HIP(ctx, hipMemset(rs.ticket, 0, sizeof(unsigned int) * 4));
    TE_HIP(ctx, hipMemset(rs.result, 0, sizeof(double2) * 64));
    return TE_OK;
}

void reduce_scratch_free(ReduceScratch& rs) {
    if (rs.partials) hipFree(rs.partials);
    if (rs.ticket) hipFree(rs.ticket);
    if (rs.result) hipFree(rs.result);
    rs = ReduceScratch{};
}

}  // namespace te
