// Device-side reduction helpers shared by the streaming kernels. All reductions are
// deterministic for a given launch shape: fixed shuffle trees inside a wavefront, fixed
// order across the waves of a block, and a last-block-done finish that sums the block
// partials in index order.
#pragma once

#include <hip/hip_runtime.h>

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

// Sums partials[0..n) (stride `stride` doubles2 between blocks) in the last block.
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

}  // namespace te
