// Device-side reduction helpers shared by the streaming kernels. All reductions are
// deterministic for a given launch shape: fixed shuffle trees inside a wavefront, fixed
// order across the waves of a block, and a last-block-done finish that sums the block
// partials in index order.
#pragma once

#include <hip/hip_runtime.h>

#include "te_internal.h"

namespace te {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over a block of kBlock (256) threads; the result is valid in every thread.
template <int NT>
__device__ __forceinline__ double2 block_sum2(double2 v, double2* sh /* >= NT/64 entries */) {
    v.x = wave_sum(v.x);
    v.y = wave_sum(v.y);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double2 r = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < NT / 64; i++) {
        r.x += sh[i].x;
        r.y += sh[i].y;
    }
    return r;
}

// Called by all threads after thread 0 stored this block's partial. Returns true in every
// thread of the block that arrives last; that block then sees every other block's partial.
__device__ __forceinline__ bool last_block_arrive(unsigned int* ticket) {
    __shared__ bool last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned int t = atomicAdd(ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}

// Sums partials[0..n) (stride `stride` doubles2 between blocks) in the calling block, in a fixed
// order (thread t takes t, t + NT, ..., then the block sum): the same result in every block.
template <int NT>
__device__ __forceinline__ double2 reduce_partials(const double2* partials, int n, int stride, double2* sh) {
    double2 acc = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < n; i += NT) {
        double2 p = partials[(size_t)i * stride];
        acc.x += p.x;
        acc.y += p.y;
    }
    return block_sum2<NT>(acc, sh);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// Logical block id remapped so that each of the 8 XCDs (which receive blocks round-robin)
// works on one contiguous eighth of the logical index space: neighbouring row ranges then
// share an L2.
__device__ __forceinline__ int xcd_block_id() {
    const int g = gridDim.x, b = blockIdx.x;
    if ((g & 7) != 0) return b;
    return (b & 7) * (g >> 3) + (b >> 3);
}

// Workgroup barrier for LDS traffic only: waits for this wave's LDS operations (lgkmcnt), not
// for its outstanding global loads, which __syncthreads() also drains. Loads issued for later
// work stay in flight across it. The memory clobber keeps the compiler from moving memory
// accesses across the barrier.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Row-block schedule of the stream kernels: XCD x (blocks b with b % 8 == x) owns the
// contiguous eighth [x nrb / 8, (x+1) nrb / 8) of the row blocks and its blocks take them
// round-robin, so they sweep that range together and the neighbours that rows a few thousand
// apart gather from x are still in the XCD's L2. Without a grid multiple of 8: round-robin.
__device__ __forceinline__ void xcd_sweep(int nrb, int& first, int& end, int& stride) {
    const int g = gridDim.x, b = blockIdx.x;
    if ((g & 7) != 0) {
        first = b;
        end = nrb;
        stride = g;
        return;
    }
    const int x = b & 7;
    first = (int)((long long)nrb * x / 8) + (b >> 3);
    end = (int)((long long)nrb * (x + 1) / 8);
    stride = g >> 3;
}

// The same at wavefront granularity, for blocks of kBlock threads: the waves of XCD x take the
// items of its eighth round-robin.
__device__ __forceinline__ void xcd_sweep_waves(int n_items, int& first, int& end, int& stride) {
    const int g = gridDim.x, b = blockIdx.x, w = threadIdx.x >> 6, wpb = blockDim.x >> 6;
    if ((g & 7) != 0) {
        first = b * wpb + w;
        end = n_items;
        stride = g * wpb;
        return;
    }
    const int x = b & 7;
    first = (int)((long long)n_items * x / 8) + (b >> 3) * wpb + w;
    end = (int)((long long)n_items * (x + 1) / 8);
    stride = (g >> 3) * wpb;
}

// ---------------------------------------------------------------------------------------
// One row block of the row-block ("CSR-stream") SpMV: rows [r0, r1) whose entries fit
// kStreamCap. Every thread issues all of its value, index and x loads (kStreamCap / kBlock of
// each) before using any, the products go to LDS (prod: kStreamCap entries), and thread t then
// sums rows r0 + t, ... in entry order, continuing acc_in[r] when given, and hands each sum to
// epi.row(r, sum). Ends with a barrier, so prod can be reused at once.
// ---------------------------------------------------------------------------------------
typedef double d2v __attribute__((ext_vector_type(2)));

// Streamed once per SpMV: loaded non-temporally so the gathered x keeps its place in L2/MALL.
__device__ __forceinline__ double2 load_stream(const double2* p) {
    const d2v t = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
    return make_double2(t.x, t.y);
}
__device__ __forceinline__ void store_stream(double2* p, double2 v) {
    d2v t;
    t.x = v.x;
    t.y = v.y;
    __builtin_nontemporal_store(t, reinterpret_cast<d2v*>(p));
}

constexpr int kNI = kStreamCap / kBlock;
static_assert(kStreamCap % kBlock == 0, "stream capacity must be a multiple of the block size");

template <class Epi>
__device__ __forceinline__ void stream_row_block(int r0, int r1, const int* __restrict__ ptr,
                                                 const int* __restrict__ col, const double2* __restrict__ val,
                                                 const double2* __restrict__ x, double2* prod, Epi& epi,
                                                 const double2* __restrict__ acc_in) {
    const int t = threadIdx.x;
    const int k0 = ptr[r0];
    const int nk = ptr[r1] - k0;
    double2 v[kNI], xv[kNI];
    int c[kNI];
#pragma unroll
    for (int i = 0; i < kNI; i++) {
        const int k = t + i * kBlock;
        if (k < nk) {
            v[i] = load_stream(&val[k0 + k]);
            c[i] = __builtin_nontemporal_load(&col[k0 + k]);
        }
    }
#pragma unroll
    for (int i = 0; i < kNI; i++) {
        const int k = t + i * kBlock;
        if (k < nk) xv[i] = x[c[i]];
    }
#pragma unroll
    for (int i = 0; i < kNI; i++) {
        const int k = t + i * kBlock;
        if (k < nk) prod[k] = cmul(v[i], xv[i]);
    }
    __syncthreads();
    for (int r = r0 + t; r < r1; r += kBlock) {
        const int a = ptr[r] - k0, e = ptr[r + 1] - k0;
        double sr = 0.0, si = 0.0;
        if (acc_in) {               // continue the row sum of the preceding column slices
            const double2 p = load_stream(&acc_in[r]);
            sr = p.x;
            si = p.y;
        }
        for (int k = a; k < e; k++) {
            sr += prod[k].x;
            si += prod[k].y;
        }
        epi.row(r, make_double2(sr, si));
    }
    __syncthreads();
}

}  // namespace te
