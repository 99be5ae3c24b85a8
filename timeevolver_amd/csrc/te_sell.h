// Device code shared by the single-instance and batched Lanczos kernels for the sliced-ELL
// value-dictionary form of a matrix (slices of kSlice rows stored slot-major, 16-bit indices into
// a table of distinct off-diagonal values, dense diagonal).
#pragma once

#include "te_internal.h"
#include "te_reduce.h"

namespace te {

// A load at a wave-uniform address of data that no kernel writes while it runs (matrix
// metadata), through the constant address space: the compiler then issues a scalar load, which
// waits on its own counter, not behind the wavefront's outstanding vector loads and stores.
__device__ __forceinline__ int scalar_load(const int* p) {
    return *(const __attribute__((address_space(4))) int*)p;
}
__device__ __forceinline__ int2 scalar_load(const int2* p) {
    const unsigned long long u = *(const __attribute__((address_space(4))) unsigned long long*)p;
    return make_int2((int)(unsigned)u, (int)(unsigned)(u >> 32));
}
__device__ __forceinline__ int4 scalar_load(const int4* p) {
    typedef int i4v __attribute__((ext_vector_type(4)));
    const i4v u = *(const __attribute__((address_space(4))) i4v*)p;
    return make_int4(u.x, u.y, u.z, u.w);
}

// Shared record q of slice s (ov: the slice's first overflow record).
__device__ __forceinline__ int2 shared_record(const SellView& m, int s, int ov, int q) {
    return q < kSellFix ? scalar_load(m.ufix + (int64_t)s * kSellFix + q) : scalar_load(m.uni + ov + (q - kSellFix));
}

// A group of kSellGroup slots: the lane's words (streamed) and the slots' base columns (the same
// for the whole wavefront: scalar loads). Slots from `valid` on lie past the slice's end: not
// loaded (padding word, base 0: column 0).
__device__ __forceinline__ void sell_load_group(const unsigned* wp, const int* bp, int valid, unsigned* w, int* b) {
#pragma unroll
    for (int q = 0; q < kSellGroup; q++) {
        w[q] = kSellPadIdx << 16;
        b[q] = 0;
        if (q < valid) {                              // the same for the whole wavefront
            w[q] = __builtin_nontemporal_load(wp + q * kSlice);
            b[q] = scalar_load(bp + q);
        }
    }
}

// Shared records q0..q0+count-1 of full slice s (the same for every row): gathers of x at
// r + offset, added in record order. Groups of G records: every gather of a group is in flight
// before any is used.
template <int NV, int G>
__device__ __forceinline__ void sell_shared_n(const SellView& m, const double2* const* x, const double2* vt, int64_t r,
                                              int s, int ov, int q0, int count, double* sr, double* si) {
    for (int t = 0; t < count; t += G) {
        double2 xv[NV][G];
        int vi[G];
#pragma unroll
        for (int q = 0; q < G; q++) {
            const int2 u = shared_record(m, s, ov, q0 + min(t + q, count - 1));   // past the end: not added
            vi[q] = u.y;
#pragma unroll
            for (int v = 0; v < NV; v++) xv[v][q] = x[v][r + u.x];
        }
#pragma unroll
        for (int q = 0; q < G; q++) {
            const bool in = t + q < count;                          // the same for the whole wavefront
            const double2 a = vt[vi[q]];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                const double2 p = cmul(a, xv[v][q]);
                sr[v] = in ? sr[v] + p.x : sr[v];
                si[v] = in ? si[v] + p.y : si[v];
            }
        }
    }
}

// Shared records a slice gathers at its start (the rest after them, in record order).
template <int NV>
struct SellShared {
    static constexpr int kIssue = NV == 1 ? 12 : NV == 2 ? 4 : 2;
    static_assert(kIssue <= kSellFix, "the issued records sit at fixed places");
};

// Slice s of a sliced-ELL value-dictionary matrix by the calling wavefront (lane = row of the
// slice), applied to NV vectors x[v] at once (one pass over the slice's words): y_r = diag_r x_r
// + sum_j vt[vidx_j] x[col_j] (diagonal first, then column order: leading shared slots, the
// per-row slots, trailing shared slots; the same sums as the row-block kernel, and for each
// vector the same as a pass of its own), handed to epi[v].row. Padding slots gather the slot's
// base column and are not added. The first KI shared records of the slice are scalar loads and
// their gathers, leading and trailing, are all issued before the per-row slots' first group (the
// rest follow in groups of 4 at their place in the sum; KI changes registers, not results). A
// group of kSellGroup per-row slots is loaded before any is used, and the next group's slots load
// while the current group's gathers are in flight. vt: the value table in LDS.
template <int NV, class Epi, int KI = SellShared<NV>::kIssue>
__device__ __forceinline__ void sell_slice_n(const SellView& m, const double2* const* x, const double2* vt, int s,
                                             int lane, Epi* epi) {
    constexpr int kIssue = KI;
    static_assert(KI >= 1 && KI <= kSellFix, "the issued records sit at fixed places");
    const int4 h = scalar_load(m.meta + s);                        // scalar loads: wave-uniform
    const int o = h.x;
    const int ls = (scalar_load(reinterpret_cast<const int*>(m.meta + s + 1)) - o) / kSlice;   // per-row slots
    const int ng = (ls + kSellGroup - 1) / kSellGroup;             // groups of slots
    const int64_t r = (int64_t)s * kSlice + lane;
    const int ov = h.y, nlo = h.z & 0xffff, nsh = h.z >> 16;       // shared records of the slice
    int2 rec[kIssue];
#pragma unroll
    for (int q = 0; q < kIssue; q++) rec[q] = make_int2(0, 0);
    if (m.ufix) {
#pragma unroll
        for (int q = 0; q < kIssue; q++) rec[q] = scalar_load(m.ufix + (int64_t)s * kSellFix + q);
    }
    double2 xr[NV], d = make_double2(0.0, 0.0);
    typename Epi::Pre pre[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        xr[v] = make_double2(0.0, 0.0);
        pre[v] = typename Epi::Pre{};
    }
    if (r < m.n) {
#pragma unroll
        for (int v = 0; v < NV; v++) xr[v] = x[v][r];
        d = m.diag_re ? make_double2(__builtin_nontemporal_load(&m.diag_re[r]), 0.0) : load_stream(&m.diag[r]);
#pragma unroll
        for (int v = 0; v < NV; v++) pre[v] = epi[v].load((int)r);
    }
    const unsigned* wp = m.word + o + lane;
    const int* bp = m.base + o / kSlice;
    unsigned w[kSellGroup];
    int b[kSellGroup];
    sell_load_group(wp, bp, ls, w, b);
    // the first kIssue shared gathers (slices with shared slots are full: every row exists)
    double2 xs[NV][kIssue];
    int vs[kIssue];
#pragma unroll
    for (int q = 0; q < kIssue; q++) {
        vs[q] = rec[q].y;
        if (q < nsh) {                                             // the same for the whole wavefront
#pragma unroll
            for (int v = 0; v < NV; v++) xs[v][q] = x[v][r + rec[q].x];
        }
    }
    double sr[NV], si[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        sr[v] = d.x * xr[v].x - d.y * xr[v].y;
        si[v] = d.x * xr[v].y + d.y * xr[v].x;
    }
    // records [q0, q1) of the issued ones, in order
    auto add_issued = [&](int q0, int q1) {
#pragma unroll
        for (int q = 0; q < kIssue; q++) {
            if (q >= q0 && q < q1) {                               // the same for the whole wavefront
                const double2 a = vt[vs[q]];
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    const double2 p = cmul(a, xs[v][q]);
                    sr[v] += p.x;
                    si[v] += p.y;
                }
            }
        }
    };
    if (nsh > 0) {
        add_issued(0, min(nlo, kIssue));
        if (nlo > kIssue) sell_shared_n<NV, 4>(m, x, vt, r, s, ov, kIssue, nlo - kIssue, sr, si);
    }
    // Group g's gathers and group g+1's word loads are in flight together (the last group is
    // peeled so that the loop body is straight-line code for the wait counters). Slots past the
    // slice end (last group) are neither loaded nor gathered.
    auto accumulate = [&](const unsigned* ww, const double2 (*xv)[kSellGroup], int g) {
#pragma unroll
        for (int q = 0; q < kSellGroup; q++) {
            const unsigned vi = ww[q] >> 16;
            const bool in = (kSellGroup * g + q < ls) & (vi != kSellPadIdx);   // selects, not branches
            const double2 a = vt[in ? vi : 0u];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                const double2 p = cmul(a, xv[v][q]);
                sr[v] = in ? sr[v] + p.x : sr[v];
                si[v] = in ? si[v] + p.y : si[v];
            }
        }
    };
    auto gather = [&](double2 (*xv)[kSellGroup], int valid) {
#pragma unroll
        for (int q = 0; q < kSellGroup; q++) {
            const int c = b[q] + (int)(w[q] & 0xffffu);
#pragma unroll
            for (int v = 0; v < NV; v++) xv[v][q] = q < valid ? x[v][c] : make_double2(0.0, 0.0);
        }
    };
    for (int g = 0; g + 1 < ng; g++) {
        double2 xv[NV][kSellGroup];
        unsigned ww[kSellGroup];
        gather(xv, kSellGroup);
#pragma unroll
        for (int q = 0; q < kSellGroup; q++) ww[q] = w[q];
        wp += kSellGroup * kSlice;
        bp += kSellGroup;
        sell_load_group(wp, bp, ls - kSellGroup * (g + 1), w, b);
        accumulate(ww, xv, g);
    }
    if (ng > 0) {
        double2 xv[NV][kSellGroup];
        gather(xv, ls - kSellGroup * (ng - 1));
        accumulate(w, xv, ng - 1);
    }
    if (nsh > nlo) {                                               // trailing shared records
        add_issued(nlo, min(nsh, kIssue));
        const int h0 = max(nlo, kIssue);
        if (nsh > h0) sell_shared_n<NV, 4>(m, x, vt, r, s, ov, h0, nsh - h0, sr, si);
    }
    if (r < m.n) {
#pragma unroll
        for (int v = 0; v < NV; v++) epi[v].row((int)r, make_double2(sr[v], si[v]), xr[v], pre[v]);
    }
}

// One vector: sell_slice_n with NV = 1 (KI shared records issued at the slice's start).
template <class Epi, int KI = SellShared<1>::kIssue>
__device__ __forceinline__ void sell_slice(const SellView& m, const double2* __restrict__ x, const double2* vt, int s,
                                           int lane, Epi& epi) {
    const double2* xs[1] = {x};
    sell_slice_n<1, Epi, KI>(m, xs, vt, s, lane, &epi);
}

// Lanczos apply epilogue on one row: w = c1 (A u_j)_r + c2 u_{j-1}[r] (non-temporal store),
// dot += conj(u_j[r]) w_r; `load` fetches u_{j-1}[r] ahead of the row's sum.
struct ApplyEpiVd {
    using Pre = double2;
    double2 c1;
    double c2;
    const double2* up;
    double2* w;
    double2 dot;
    __device__ __forceinline__ Pre load(int r) const {
        return up ? load_stream(&up[r]) : make_double2(0.0, 0.0);
    }
    __device__ __forceinline__ void row(int r, double2 s, double2 u, Pre p) {
        double2 wr = cmul(c1, s);
        if (up) {
            wr.x += c2 * p.x;
            wr.y += c2 * p.y;
        }
        store_stream(&w[r], wr);
        dot.x += u.x * wr.x + u.y * wr.y;
        dot.y += u.x * wr.y - u.y * wr.x;
    }
};

}  // namespace te
