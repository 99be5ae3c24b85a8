// Batched Krylov time evolution: many independent instances of the plain method (each with its
// own Hamiltonian, dimension, initial state and diagonal observables, all with one parameter
// set) advanced together. It has no counterpart in the reference, whose cost for such a batch
// is one krylovTimeEvolver::timeEvolve (core/krylovTimeEvolver.cpp:261-512) per instance.
//
// Every instance keeps the reference's algorithm unchanged: the same Lanczos recurrence, step
// controller, sample schedule and status rules as the single-instance evolver. What changes is
// how the work reaches the device:
//  * the Krylov bases of all instances are packed into one array of (m_max + 1) columns, an
//    instance owning rows [off_i, off_i + D_i) of every column;
//  * one launch per kernel kind covers every instance: the Lanczos apply walks SpMV units (a run
//    of sliced-ELL slices of one instance when every instance has its value-dictionary copy,
//    vectors then in each instance's own locality order; else up to kBatchRB CSR row blocks),
//    the project/combine passes walk vector units (up to vrows rows of one instance); each unit
//    first reads its instance's flags and skips the instance when it is finished, broken down,
//    or past its Krylov dimension;
//  * reductions are per instance, deterministic and independent of the launch shape: a
//    wavefront writes one partial per sliced-ELL slice (apply) or per kBatchChunk-row chunk
//    (projection, observable sums), and batch_finish_kernel / batch_obs_finish_kernel (one
//    block per instance, launches of their own) sum the instance's partials in index order, so
//    unit sizes and grid sizes (which the tuner picks) leave every result bit unchanged;
//  * the unit sizes and blocks per CU are measured on the first run (a few candidates, one
//    Lanczos iteration each) unless the caller fixes them, and cached per instance-size key;
//  * the step controller runs one block per instance (launch_batch_step_control);
//  * one round (m_max Lanczos iterations, controller, combine, observable sums, live count) is
//    captured once as a HIP graph and replayed until no instance is live. The graph holds
//    kernel nodes only.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "te_internal.h"
#include "te_krylov.h"
#include "te_reduce.h"
#include "te_sell.h"

namespace te {

constexpr int kBatchRB = 4;          // row blocks (kStreamChunk entries + rows each) per CSR SpMV unit
// Default slices (kSlice rows each) per sliced-ELL SpMV unit: the batch's slices per block of the
// grid, clamped to [16, 32]. A unit may restage its instance's value table; on 512 ring
// instances with L in {6..10} (181k slices, 4096-row vector units) the batched run took
// 0.1473 / 0.1462 / 0.1470 / 0.1489 s with 16 / 24 / 32 / 48-slice units, 0.1601 s with 88.
constexpr int kBatchSlicesMin = 16, kBatchSlicesMax = 32;
// Default rows per vector unit (projection, combine): 2048 / 4096 rows gave 0.1467 / 0.1470 s
// with 32-slice units; 4096 / 8192 / 16384 rows 0.1601 / 0.1628 / 0.1667 s with 88-slice units.
constexpr int kBatchVecRows = 4096;
constexpr int kBatchBlocksPerCU = 8; // default blocks per CU of the streaming launches
// Rows per reduction chunk of the vector passes: one wavefront's share (8 rows a lane), whose
// partial is one entry of the instance's partials. Vector units are whole chunks.
constexpr int kBatchChunk = 8 * 64;
// Slices per reduction group of the sliced-ELL apply: a wavefront takes the group's slices one
// after the other and writes the group's <u_j, w> as one partial. SpMV units are whole groups.
constexpr int kBatchGroup = 4;
constexpr int kBatchNV = 4;          // instances of one shared matrix served by one pass over its words
constexpr int kBatchSC = 8;          // samples per pass of the combine kernel
constexpr int kBatchObsPass = 4;     // diagonal observables per combine pass over U (registers)
constexpr int kBatchMaxObs = 64;     // diagonal observables per instance
constexpr int kBatchTraceCap = 4096; // recorded step sizes per instance

struct BatchInst {
    int64_t off;            // first row of the instance in the packed vectors
    int n;                  // dimension
    int nsu, su0;           // CSR SpMV units [su0, su0 + nsu) (row-block path; fixed size)
    int nsl, sl0;           // apply partials [sl0, sl0 + nsl): one per group of kBatchGroup slices
                            // (sliced ELL) or per CSR unit
    int nch, ch0;           // kBatchChunk-row chunks [ch0, ch0 + nch): projection and observable partials
    const int* ptr;         // the instance's CSR (rows and columns local to the instance)
    const int* col;
    const double2* val;
    const int* rb;          // its row blocks (kStreamChunk merge-path chunks)
    SellView sell;          // its sliced-ELL value-dictionary copy (when the batch uses it)
};

struct BatchDev {
    const BatchInst* inst;
    const int4* sunit;      // (instance, first row block, end row block, 0), or slices for sliced ELL
    int n_sunits;
    const int4* gunit;      // shared-matrix units (first instance, first slice, end slice, g + 1)
    int n_gunits;
    const int4* vgroup;     // shared-matrix groups: up to kBatchNV instances (-1 past the group)
    const int4* vunit;      // (instance, first row, end row, 0), rows local to the instance; rows
    int n_vunits;           // start on multiples of kBatchChunk
    // A round's plan (batch_plan_kernel): the units of its live instances, in table order and
    // split into pieces when few are live; counts {sliced ELL, shared matrix, vector}.
    int4* lsunit;
    int4* lgunit;
    int4* lvunit;
    int* n_plan;
    double2* U;             // packed Krylov bases, column-major, ld rows per column
    int64_t ld;
    double2 alpha;          // -i expFactor
    double tolRate;
    KrylovScalars* ks;      // per instance
    StepState* st;
    BatchCtl* bc;
    double2* spart;         // per slice (sliced ELL) or per CSR SpMV unit
    double2* vpart;         // per chunk
    const double2* wcoef;   // per instance, kMaxKrylov
    const double2* Y;       // per instance, n_samples x kMaxKrylov
    long long n_samples;
    const double* diag;     // nobs x ld: Re of the diagonal observables, packed like the vectors
    int nobs;
    double* opart;          // chunks x nobs x n_samples
    double* obs;            // n_inst x nobs x n_samples
    double2* sampled;       // packed last sampled states
    double* nrm;            // per instance |sampled|^2 (final pass)
    int n_inst;
    int* n_live;
};

// Whether instance i takes part in Lanczos iteration j of this round (uniform in a block).
__device__ __forceinline__ bool lanczos_active(const BatchDev& a, int i, int j) {
    return a.bc[i].live && !a.ks[i].broken && j < a.st[i].m;
}

struct BatchApplyEpi {
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

// Lanczos apply for iteration j of every active instance: w = (alpha s_j) A u_j +
// (beta_{j-1} s_{j-1}) u_{j-1}, a_j = s_j <u_j, w> (the single-instance apply kernel's
// arithmetic, per row block).
__global__ __launch_bounds__(kBlock) void batch_apply_kernel(BatchDev a, int j) {
    __shared__ double2 prod[kStreamCap];
    __shared__ double2 sh[kBlock / 64];
    int u0, u1, us;
    xcd_sweep(a.n_sunits, u0, u1, us);
    for (int u = u0; u < u1; u += us) {
        const int4 su = a.sunit[u];
        const int i = su.x;
        if (!lanczos_active(a, i, j)) continue;
        const BatchInst I = a.inst[i];
        KrylovScalars* ks = a.ks + i;
        const double sj = ks->s[j];
        BatchApplyEpi epi;
        epi.c1 = make_double2(a.alpha.x * sj, a.alpha.y * sj);
        epi.c2 = j > 0 ? ks->beta[j - 1] * ks->s[j - 1] : 0.0;
        epi.up = j > 0 ? a.U + (size_t)(j - 1) * a.ld + I.off : nullptr;
        epi.uj = a.U + (size_t)j * a.ld + I.off;
        epi.w = a.U + (size_t)(j + 1) * a.ld + I.off;
        epi.dot = make_double2(0.0, 0.0);
        for (int b = su.y; b < su.z; b++)
            stream_row_block(I.rb[b], I.rb[b + 1], I.ptr, I.col, I.val, epi.uj, prod, epi, nullptr);
        const double2 bs = block_sum2<kBlock>(epi.dot, sh);
        if (threadIdx.x == 0) a.spart[I.sl0 + (u - I.su0)] = bs;      // the unit's partial
    }
}

// The same for instances held as sliced-ELL value-dictionary copies (vectors in each instance's
// locality order): a unit is a run of slices of one instance, its wavefronts take the
// slices in turn, and the instance's value table is staged in LDS when the unit's instance
// differs from the block's previous one. A wavefront takes a group of kBatchGroup consecutive
// slices; the group's <u_j, w> goes to the group's partial (a wavefront sum), whatever unit or
// block computed it.
// The apply epilogue of one instance in a shared-matrix pass; an instance that takes no part
// in the iteration (finished, broken down, or past its Krylov dimension) is carried along with
// its loads and stores switched off.
struct ApplyEpiOpt {
    using Pre = double2;
    ApplyEpiVd e;
    bool on;
    __device__ __forceinline__ Pre load(int r) const { return on ? e.load(r) : make_double2(0.0, 0.0); }
    __device__ __forceinline__ void row(int r, double2 sum, double2 u, Pre p) {
        if (on) e.row(r, sum, u, p);
    }
};

__device__ __forceinline__ ApplyEpiVd batch_apply_epi(const BatchDev& a, int i, int j) {
    const BatchInst& I = a.inst[i];
    const KrylovScalars* ks = a.ks + i;
    const double sj = ks->s[j];
    ApplyEpiVd epi;
    epi.c1 = make_double2(a.alpha.x * sj, a.alpha.y * sj);
    epi.c2 = j > 0 ? ks->beta[j - 1] * ks->s[j - 1] : 0.0;
    epi.up = j > 0 ? a.U + (size_t)(j - 1) * a.ld + I.off : nullptr;
    epi.w = a.U + (size_t)(j + 1) * a.ld + I.off;
    epi.dot = make_double2(0.0, 0.0);
    return epi;
}

// A unit of shared-matrix group g: NV instances (slots of the group past its size, or
// instances not in this iteration, switched off) in one pass over the matrix words; each
// instance's rows, sums and group partials are those of a pass of its own.
template <int NV>
__device__ void batch_apply_group(const BatchDev& a, int j, int4 su, const double2* vt, int lane, int wv) {
    const int4 gi = a.vgroup[su.w - 1];
    const int ids[4] = {gi.x, gi.y, gi.z, gi.w};
    const SellView m = a.inst[gi.x].sell;
    ApplyEpiOpt epi[NV];
    const double2* xs[NV];
    int sl0[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const int i = ids[v];
        const bool on = i >= 0 && lanczos_active(a, i, j);
        const int ii = i >= 0 ? i : gi.x;
        epi[v].e = batch_apply_epi(a, ii, j);
        epi[v].on = on;
        xs[v] = a.U + (size_t)j * a.ld + a.inst[ii].off;
        sl0[v] = a.inst[ii].sl0;
    }
    const int g1 = (su.z + kBatchGroup - 1) / kBatchGroup;
    for (int g = su.y / kBatchGroup + wv; g < g1; g += kBlock / 64) {
        const int gw = __builtin_amdgcn_readfirstlane(g);
        const int se = min((gw + 1) * kBatchGroup, su.z);
#pragma unroll
        for (int v = 0; v < NV; v++) epi[v].e.dot = make_double2(0.0, 0.0);
        for (int sl = gw * kBatchGroup; sl < se; sl++) sell_slice_n<NV>(m, xs, vt, sl, lane, epi);
#pragma unroll
        for (int v = 0; v < NV; v++) {
            const double dx = wave_sum(epi[v].e.dot.x), dy = wave_sum(epi[v].e.dot.y);
            if (lane == 0 && epi[v].on) a.spart[sl0[v] + gw] = make_double2(dx, dy);
        }
    }
}

// Units of shared-matrix groups (a launch of their own: the four-vector pass needs about twice
// the registers of the single one).
__global__ __launch_bounds__(kBlock) void batch_apply_group_kernel(BatchDev a, int j) {
    __shared__ double2 vt[kVdMaxValues];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const double2* staged = nullptr;
    int u0, u1, us;
    xcd_sweep(a.n_plan[1], u0, u1, us);
    for (int u = u0; u < u1; u += us) {
        const int4 su = a.lgunit[u];
        const int4 gi = a.vgroup[su.w - 1];
        const bool any = lanczos_active(a, gi.x, j) || (gi.y >= 0 && lanczos_active(a, gi.y, j)) ||
                         (gi.z >= 0 && lanczos_active(a, gi.z, j)) || (gi.w >= 0 && lanczos_active(a, gi.w, j));
        if (!any) continue;
        const SellView& mg = a.inst[gi.x].sell;
        if (mg.vtab != staged) {
            __syncthreads();                             // every wavefront is done with the old table
            for (int q = threadIdx.x; q < mg.n_values; q += kBlock) vt[q] = mg.vtab[q];
            __syncthreads();
            staged = mg.vtab;
        }
        if (gi.z >= 0) batch_apply_group<4>(a, j, su, vt, lane, wv);
        else batch_apply_group<2>(a, j, su, vt, lane, wv);
    }
}

// Four wavefronts per SIMD (at most 128 registers): a slice is a chain of dependent loads, so
// more wavefronts hide more of it. With the 12 shared-record gathers the single-instance apply
// issues at a slice's start the compiler takes 143 registers (3 per SIMD); the batch's instances
// are small rings (L <= 10: at most 8 shared records a slice, 2.9 on average at L = 10), so a
// slice issues kBatchIssue of them up front and the rest in groups of 4 in their order (the same
// sums): 111 registers. On 512 ring instances with L in {6..10}: apply 97.8 ms a run at 143
// registers, 83.9 at 128 (12 issued, 8 values spilled), 78.7 at 111 (4 issued); 2 issued at 5
// waves per SIMD 78.5, 6 issued 79.2.
constexpr int kBatchIssue = 4;
__global__ __launch_bounds__(kBlock, 4) void batch_apply_sell_kernel(BatchDev a, int j) {
    __shared__ double2 vt[kVdMaxValues];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const double2* staged = nullptr;
    int u0, u1, us;
    xcd_sweep(a.n_plan[0], u0, u1, us);
    for (int u = u0; u < u1; u += us) {
        const int4 su = a.lsunit[u];
        const int i = su.x;
        if (!lanczos_active(a, i, j)) continue;
        const BatchInst& I = a.inst[i];
        const SellView m = I.sell;
        if (m.vtab != staged) {
            __syncthreads();                             // every wavefront is done with the old table
            for (int q = threadIdx.x; q < m.n_values; q += kBlock) vt[q] = m.vtab[q];
            __syncthreads();
            staged = m.vtab;
        }
        KrylovScalars* ks = a.ks + i;
        const double sj = ks->s[j];
        ApplyEpiVd epi;
        epi.c1 = make_double2(a.alpha.x * sj, a.alpha.y * sj);
        epi.c2 = j > 0 ? ks->beta[j - 1] * ks->s[j - 1] : 0.0;
        epi.up = j > 0 ? a.U + (size_t)(j - 1) * a.ld + I.off : nullptr;
        epi.w = a.U + (size_t)(j + 1) * a.ld + I.off;
        epi.dot = make_double2(0.0, 0.0);
        const double2* x = a.U + (size_t)j * a.ld + I.off;
        const int g1 = (su.z + kBatchGroup - 1) / kBatchGroup;
        for (int g = su.y / kBatchGroup + wv; g < g1; g += kBlock / 64) {
            const int gw = __builtin_amdgcn_readfirstlane(g);
            const int se = min((gw + 1) * kBatchGroup, su.z);
            epi.dot = make_double2(0.0, 0.0);
            for (int sl = gw * kBatchGroup; sl < se; sl++) sell_slice<ApplyEpiVd, kBatchIssue>(m, x, vt, sl, lane, epi);
            const double dx = wave_sum(epi.dot.x), dy = wave_sum(epi.dot.y);
            if (lane == 0) a.spart[I.sl0 + gw] = make_double2(dx, dy);   // the group's partial
        }
    }
}

// Lanczos project for iteration j: w' = w - (a_j s_j) u_j, beta_j = |w'|, stored in place
// unless j + 1 is the instance's Krylov dimension; breakdown when beta_j < tolRate
// (core/krylovTimeEvolver.cpp:551-567). A wavefront takes a chunk of kBatchChunk rows (lane l:
// rows l, l + 64, ..., in order) and writes the chunk's |w'|^2 as its partial.
// u_j and w are read and w' written non-temporally (on 512 ring instances with L in {6..10} the
// projection took 22.1-22.2 ms a run against 24.1 with ordinary loads and stores, and the next
// apply, which gathers w', was no slower).
__global__ __launch_bounds__(kBlock) void batch_project_kernel(BatchDev a, int j) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int u0, u1, us;
    xcd_sweep(a.n_plan[2], u0, u1, us);
    for (int u = u0; u < u1; u += us) {
        const int4 vu = a.lvunit[u];
        const int i = vu.x;
        if (!lanczos_active(a, i, j)) continue;
        const BatchInst& I = a.inst[i];
        KrylovScalars* ks = a.ks + i;
        const int m = a.st[i].m;
        const double sj = ks->s[j];
        const double2 aj = ks->a[j];
        const double2 c3 = make_double2(aj.x * sj, aj.y * sj);
        const double2* __restrict__ uj = a.U + (size_t)j * a.ld + I.off;
        double2* __restrict__ w = a.U + (size_t)(j + 1) * a.ld + I.off;
        const bool store = j + 1 < m;
        const int cb = vu.y / kBatchChunk, ce = (vu.z + kBatchChunk - 1) / kBatchChunk;
        for (int c = cb + wv; c < ce; c += kBlock / 64) {
            constexpr int kR = kBatchChunk / 64;
            const int r0 = c * kBatchChunk + lane;
            double2 uu[kR], wr[kR];
#pragma unroll
            for (int q = 0; q < kR; q++) {
                const int r = r0 + 64 * q;
                uu[q] = make_double2(0.0, 0.0);
                wr[q] = make_double2(0.0, 0.0);
                if (r < vu.z) {                   // streamed: read and written once a pass
                    uu[q] = load_stream(&uj[r]);
                    wr[q] = load_stream(&w[r]);
                }
            }
            double nrm = 0.0;
#pragma unroll
            for (int q = 0; q < kR; q++) {
                const int r = r0 + 64 * q;
                wr[q].x -= c3.x * uu[q].x - c3.y * uu[q].y;
                wr[q].y -= c3.x * uu[q].y + c3.y * uu[q].x;
                if (r < vu.z) {
                    if (store) store_stream(&w[r], wr[q]);
                    nrm += wr[q].x * wr[q].x + wr[q].y * wr[q].y;
                }
            }
            nrm = wave_sum(nrm);
            if (lane == 0) a.vpart[I.ch0 + c] = make_double2(nrm, 0.0);
        }
    }
}

// One block per instance after an apply (PROJECT false: alpha_j = s_j <u_j, w>) or a projection
// (PROJECT true: beta_j = |w'|, stored with s_{j+1} = 1 / beta_j, or h when j + 1 is the Krylov
// dimension; breakdown when beta_j < tolRate, core/krylovTimeEvolver.cpp:551-567) of iteration j:
// the instance's slice (or CSR unit) or chunk partials summed in index order. In a launch of its own: a reduction by the
// instance's last unit inside the apply or projection costs every unit an agent-scope release,
// which on this multi-XCD part writes back the XCD's L2 (per-unit tickets made the projection
// of 512 instances about 1.7x slower in a probe of its shape).
template <bool PROJECT>
__global__ __launch_bounds__(kBlock) void batch_finish_kernel(BatchDev a, int j) {
    __shared__ double2 sh[kBlock / 64];
    const int i = blockIdx.x;
    if (i >= a.n_inst || !lanczos_active(a, i, j)) return;
    const BatchInst& I = a.inst[i];
    KrylovScalars* ks = a.ks + i;
    const double2 tot = PROJECT ? reduce_partials<kBlock>(a.vpart + I.ch0, I.nch, 1, sh)
                                : reduce_partials<kBlock>(a.spart + I.sl0, I.nsl, 1, sh);
    if (threadIdx.x != 0) return;
    if (!PROJECT) {
        const double sj = ks->s[j];
        ks->a[j] = make_double2(sj * tot.x, sj * tot.y);
        return;
    }
    const int m = a.st[i].m;
    const double beta = sqrt(tot.x);
    if (beta < a.tolRate) {
        ks->broken = 1;
        ks->m_hbd = j + 1;
        ks->h = beta;
    } else if (j + 1 == m) {
        ks->h = beta;
    } else {
        ks->beta[j] = beta;
        ks->s[j + 1] = 1.0 / beta;
    }
}

// After the controller: for every instance that stepped, currentVec = U w written over column
// 0 when the step asks for it (:382), the step's samples U y_s (:409-421), the last of them
// kept as the sampled state, and per-chunk partial sums of the diagonal observables
// sum_r Re(O_rr) |psi_s(r)|^2 (a wavefront per kBatchChunk-row chunk, lane l: rows l, l + 64,
// ...). Samples go kBatchSC per pass over U; the pass that covers the last sample is the one
// that overwrites column 0, after every pass has read it. Observables go NOB (at most
// kBatchObsPass) a launch, from obs0 on; only the launch of the last group (last_group) writes
// column 0 and the sampled state, so every launch reads the step's own basis.
template <int NOB>
__global__ __launch_bounds__(kBlock) void batch_combine_kernel(BatchDev a, int obs0, int last_group) {
    constexpr int kBatchBurst = 8;
    __shared__ double2 cf[(1 + kBatchSC) * kMaxKrylov];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int u0, u1, us;
    xcd_sweep(a.n_plan[2], u0, u1, us);
    for (int u = u0; u < u1; u += us) {
        const int4 vu = a.lvunit[u];
        const int i = vu.x;
        if (!a.bc[i].stepped) continue;
        const int m = a.bc[i].cm;
        const StepState* st = a.st + i;
        const int S = st->S;
        const long long y0 = st->sample_base;
        const int write_w = st->write_w;
        const BatchInst& I = a.inst[i];
        const int64_t off = I.off;
        const double2* Yi = a.Y + (size_t)i * a.n_samples * kMaxKrylov;
        const double2* wc = a.wcoef + (size_t)i * kMaxKrylov;
        const int cb = vu.y / kBatchChunk, ce = (vu.z + kBatchChunk - 1) / kBatchChunk;
        for (int c0 = 0; c0 < S || (c0 == 0 && write_w); c0 += kBatchSC) {
            const int sc = min(kBatchSC, S - c0);      // samples of this pass (may be <= 0)
            const bool lastpass = c0 + kBatchSC >= S;
            const bool ww = lastpass && write_w && last_group;
            for (int q = threadIdx.x; q < (1 + kBatchSC) * kMaxKrylov; q += kBlock) {
                const int c = q / kMaxKrylov, jj = q % kMaxKrylov;
                double2 v = make_double2(0.0, 0.0);
                if (jj < m) {
                    if (c == 0) {
                        if (ww) v = wc[jj];
                    } else if (c - 1 < sc) {
                        v = Yi[(size_t)(y0 + c0 + c - 1) * kMaxKrylov + jj];
                    }
                }
                cf[q] = v;
            }
            __syncthreads();
            for (int ch = cb + wv; ch < ce; ch += kBlock / 64) {
                double acc[NOB > 0 ? NOB * kBatchSC : 1];
#pragma unroll
                for (int q = 0; q < (NOB > 0 ? NOB * kBatchSC : 1); q++) acc[q] = 0.0;
                for (int qr = 0; qr < kBatchChunk / 64; qr++) {
                    const int row = ch * kBatchChunk + lane + 64 * qr;
                    if (row >= vu.z) break;
                    const size_t gr = (size_t)off + row;
                    double2 aw = make_double2(0.0, 0.0);
                    double2 as[kBatchSC];
#pragma unroll
                    for (int s = 0; s < kBatchSC; s++) as[s] = make_double2(0.0, 0.0);
                    auto add = [&](int jj, double2 uv) {
                        if (ww) {
                            const double2 t = cmul(uv, cf[jj]);
                            aw.x += t.x;
                            aw.y += t.y;
                        }
#pragma unroll
                        for (int s = 0; s < kBatchSC; s++) {
                            if (s < sc) {
                                const double2 t = cmul(uv, cf[(1 + s) * kMaxKrylov + jj]);
                                as[s].x += t.x;
                                as[s].y += t.y;
                            }
                        }
                    };
                    // kBatchBurst basis columns in flight per lane (sums still in order of jj), read
                    // non-temporally: a column is read once a pass (combine 11.4-11.5 -> 10.7-10.9
                    // ms a run of the 512-instance batch)
                    int jj = 0;
                    for (; jj + kBatchBurst <= m; jj += kBatchBurst) {
                        double2 uv[kBatchBurst];
#pragma unroll
                        for (int q = 0; q < kBatchBurst; q++) uv[q] = load_stream(&a.U[(size_t)(jj + q) * a.ld + gr]);
#pragma unroll
                        for (int q = 0; q < kBatchBurst; q++) add(jj + q, uv[q]);
                    }
                    for (; jj < m; jj++) add(jj, load_stream(&a.U[(size_t)jj * a.ld + gr]));
                    if (ww) a.U[gr] = aw;
                    if (last_group && lastpass && sc > 0) {
                        double2 v = as[0];
#pragma unroll
                        for (int s = 1; s < kBatchSC; s++)
                            if (s == sc - 1) v = as[s];
                        a.sampled[gr] = v;
                    }
                    double dg[NOB > 0 ? NOB : 1];              // each observable's diagonal read once a row
#pragma unroll
                    for (int o = 0; o < NOB; o++)
                        dg[o] = obs0 + o < a.nobs ? a.diag[(size_t)(obs0 + o) * a.ld + gr] : 0.0;
#pragma unroll
                    for (int s = 0; s < kBatchSC; s++) {
                        const double pr = as[s].x * as[s].x + as[s].y * as[s].y;
#pragma unroll
                        for (int o = 0; o < NOB; o++)
                            if (obs0 + o < a.nobs && s < sc) acc[o * kBatchSC + s] += dg[o] * pr;
                    }
                }
#pragma unroll
                for (int q = 0; q < NOB * kBatchSC; q++) {
                    const int o = q / kBatchSC, s = q % kBatchSC;
                    if (obs0 + o < a.nobs && s < sc) {   // uniform in the wavefront
                        const double t = wave_sum(acc[q]);
                        if (lane == 0)
                            a.opart[((size_t)(I.ch0 + ch) * a.nobs + obs0 + o) * a.n_samples + y0 + c0 + s] = t;
                    }
                }
            }
            __syncthreads();                     // cf is reloaded by the next pass
        }
    }
}

// One block per instance that stepped: its observable values for the step's samples, summed
// over its chunks in index order.
__global__ __launch_bounds__(kBlock) void batch_obs_finish_kernel(BatchDev a) {
    const int i = blockIdx.x;
    if (!a.bc[i].stepped || a.nobs == 0) return;
    const int S = a.st[i].S;
    const long long y0 = a.st[i].sample_base;
    const BatchInst I = a.inst[i];
    for (int p = threadIdx.x; p < a.nobs * S; p += kBlock) {
        const int o = p / S, s = p % S;
        double t = 0.0;
        for (int q = 0; q < I.nch; q++) t += a.opart[((size_t)(I.ch0 + q) * a.nobs + o) * a.n_samples + y0 + s];
        a.obs[((size_t)i * a.nobs + o) * a.n_samples + y0 + s] = t;
    }
}

// A round's plan: the unit tables filtered to the instances live at the round's start (vector
// units also to those whose step the combine may still take), in table order. When fewer units
// than `target` are left, each is cut f = ceil(target / left) ways into pieces of whole
// reduction groups (kBatchGroup slices) or chunks (kBatchChunk rows), at least one of them a
// wavefront, so that the launches of a shrinking live set still spread over every XCD (each
// takes an eighth of the planned list) and CU. Partials are per group and per chunk: the plan
// changes no result bit. One block; the sliced-ELL list only on the sliced-ELL path.
constexpr int kPlanThreads = 1024;
constexpr int kPlanUnitsPerCU = 8;      // the target: planned units per CU

// Exclusive prefix sum of v over the block's threads in thread order; *total gets the sum.
__device__ int plan_scan(int v, int* sh, int* total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < kPlanThreads; d <<= 1) {
        const int add = t >= d ? sh[t - d] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const int incl = sh[t];
    *total = sh[kPlanThreads - 1];
    __syncthreads();
    return incl - v;
}

// Piece length for a unit of len items cut f ways: whole blocks of `align` items, at least
// min_len items (one piece when the unit is shorter).
__device__ __forceinline__ int plan_piece(int len, int f, int min_len, int align) {
    const int p = ((len + f - 1) / f + align - 1) / align * align;
    return max(p, min_len);
}

// The pieces of the live units of in[0, n) to out (capacity: one piece per align items of the
// table), their number to *count. Thread t takes a contiguous run of units.
template <class Live>
__device__ void plan_list(const int4* in, int n, int4* out, int* count, int target, int min_len, int align,
                          Live live, int* sh) {
    const int per = (n + kPlanThreads - 1) / kPlanThreads;
    const int u0 = min(n, (int)threadIdx.x * per), u1 = min(n, u0 + per);
    int c = 0;
    for (int u = u0; u < u1; u++) c += live(in[u]) ? 1 : 0;
    int left = 0;
    plan_scan(c, sh, &left);
    const int f = (left == 0 || left >= target) ? 1 : (target + left - 1) / left;
    int np = 0;
    for (int u = u0; u < u1; u++) {
        const int4 e = in[u];
        if (!live(e)) continue;
        const int len = e.z - e.y, pl = plan_piece(len, f, min_len, align);
        np += (len + pl - 1) / pl;
    }
    int total = 0;
    int o = plan_scan(np, sh, &total);
    for (int u = u0; u < u1; u++) {
        const int4 e = in[u];
        if (!live(e)) continue;
        const int pl = plan_piece(e.z - e.y, f, min_len, align);
        for (int k = e.y; k < e.z; k += pl) out[o++] = make_int4(e.x, k, min(k + pl, e.z), e.w);
    }
    if (threadIdx.x == 0) *count = total;
}

__global__ __launch_bounds__(kPlanThreads) void batch_plan_kernel(BatchDev a, int target, int sell) {
    __shared__ int sh[kPlanThreads];
    const BatchCtl* bc = a.bc;
    const int4* vg = a.vgroup;
    constexpr int kWaves = kBlock / 64;
    plan_list(a.sunit, sell ? a.n_sunits : 0, a.lsunit, a.n_plan + 0, target, kWaves * kBatchGroup, kBatchGroup,
              [&](int4 e) { return bc[e.x].live != 0; }, sh);
    plan_list(a.gunit, a.n_gunits, a.lgunit, a.n_plan + 1, target, kWaves * kBatchGroup, kBatchGroup,
              [&](int4 e) {
                  const int4 g = vg[e.w - 1];
                  return bc[g.x].live != 0 || (g.y >= 0 && bc[g.y].live != 0) || (g.z >= 0 && bc[g.z].live != 0) ||
                         (g.w >= 0 && bc[g.w].live != 0);
              }, sh);
    plan_list(a.vunit, a.n_vunits, a.lvunit, a.n_plan + 2, target, kWaves * kBatchChunk, kBatchChunk,
              [&](int4 e) { return bc[e.x].live != 0 || bc[e.x].stepped != 0; }, sh);
}

// Number of live instances, read by the host after each round.
__global__ __launch_bounds__(kBlock) void batch_count_kernel(BatchDev a) {
    __shared__ double2 sh[kBlock / 64];
    double c = 0.0;
    for (int i = threadIdx.x; i < a.n_inst; i += kBlock) c += a.bc[i].live ? 1.0 : 0.0;
    const double2 t = block_sum2<kBlock>(make_double2(c, 0.0), sh);
    if (threadIdx.x == 0) *a.n_live = (int)t.x;
}

// |sampled_i|^2 per instance (one block each, rows in order) for the final norm check.
__global__ __launch_bounds__(kBlock) void batch_norm_kernel(BatchDev a) {
    __shared__ double2 sh[kBlock / 64];
    const int i = blockIdx.x;
    const BatchInst I = a.inst[i];
    double s = 0.0;
    for (int r = threadIdx.x; r < I.n; r += kBlock) {
        const double2 v = a.sampled[(size_t)I.off + r];
        s += v.x * v.x + v.y * v.y;
    }
    const double2 t = block_sum2<kBlock>(make_double2(s, 0.0), sh);
    if (threadIdx.x == 0) a.nrm[i] = t.x;
}

}  // namespace te

using namespace te;

struct te_batch {
    te_ctx* ctx = nullptr;
    te_evolve_params P{};
    int n_inst = 0, nobs = 0, m_max = 0;
    long long n_samples = 0;
    int64_t ld = 0;
    double tolRate = 0;
    std::vector<int64_t> off;
    std::vector<int> dim, m0;
    std::vector<double> numErrEst;
    BatchDev dev{};
    int grid_s = 8, grid_v = 8;
    bool sell = false;                 // instances held as sliced-ELL copies (locality order)
    std::vector<int> n_slices, n_rblocks;   // per instance: sliced-ELL slices, CSR row blocks
    int64_t total_slices = 0, total_groups = 0, total_chunks = 0;
    // launch shape: slices per SpMV unit, rows per vector unit, blocks per CU of the SpMV and
    // the vector launches; source 0 default, 1 measured by the tuner, 2 the tuner's cache, 3 set
    // by the caller. tune: measure on the next run.
    int unit_slices = 0, vec_rows = 0, spmv_bpc = 0, vec_bpc = 0, shape_source = 0;
    bool tune = true;
    double tune_ms = 0.0;
    uint64_t shape_key = 0;
    std::vector<const int*> perm, iperm;
    // shared-matrix groups (sliced ELL): instances with the same matrix, up to kBatchNV a group
    // (sizes 4 and 2; a single instance keeps its own units); group_of[i] = g, or -1
    std::vector<int4> groups;
    std::vector<int> group_of;
    int4* vgroup = nullptr;
    double2* tmp = nullptr;            // packed staging vector (reordering)
    // device buffers
    BatchInst* inst = nullptr;
    int4 *sunit = nullptr, *gunit = nullptr, *vunit = nullptr;
    int4 *lsunit = nullptr, *lgunit = nullptr, *lvunit = nullptr;   // a round's plan
    int* n_plan = nullptr;
    int grid_g = 8;
    double2 *U = nullptr, *psi0 = nullptr, *sampled = nullptr, *wcoef = nullptr, *Y = nullptr;
    double2 *spart = nullptr, *vpart = nullptr;
    KrylovScalars* ks = nullptr;
    StepState* st = nullptr;
    BatchCtl* bc = nullptr;
    StepConsts* consts = nullptr;
    double *diag = nullptr, *opart = nullptr, *obs = nullptr, *nrm = nullptr, *trace = nullptr;
    int* n_live = nullptr;
    int* n_live_host = nullptr;     // pinned
    DevQuad quad;
    // one round as a graph
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long long rounds = 0;
    double round_ms = 0;
    // optional timing by launch kind (HIP event pairs on the context stream); while it is on,
    // rounds are launched kernel by kernel (no graph replay; the results are the same)
    int prof_kinds = 0;
    std::vector<hipEvent_t> prof_ev[4];
    size_t prof_used[4] = {0, 0, 0, 0};
    double prof_ms[4] = {0, 0, 0, 0};
    size_t prof_n[4] = {0, 0, 0, 0};
    std::vector<double> prof_each[4];  // the last run's launches, in launch order
};

namespace te {

static void batch_free(te_batch* b) {
    if (!b) return;
    hipSetDevice(b->ctx->device);
    for (auto& v : b->prof_ev)
        for (hipEvent_t e : v) hipEventDestroy(e);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    void* ptrs[] = {b->inst, b->sunit,   b->gunit, b->vgroup, b->vunit,  b->U,  b->psi0, b->sampled, b->wcoef, b->Y, b->tmp,
                    b->spart, b->vpart,  b->ks, b->st, b->bc, b->consts,
                    b->diag, b->opart,   b->obs,    b->nrm, b->trace, b->n_live, b->lsunit, b->lgunit,
                    b->lvunit, b->n_plan};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (b->n_live_host) hipHostFree(b->n_live_host);
    free_quadrature(b->quad);
    delete b;
}

template <class T>
static int balloc(te_ctx* ctx, T** p, size_t count) {
    TE_HIP(ctx, hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1)));
    return TE_OK;
}

constexpr size_t kBatchProfPairs = 8192;   // event pairs per launch kind

// Launch kinds of the timing: 0 apply (with its finish launch), 1 projection (with its finish
// launch), 2 step controller, 3 combine (with the observable sums and the live count).
static void prof_begin(te_batch* b, int kind, hipStream_t s) {
    if (((b->prof_kinds >> kind) & 1) && b->prof_used[kind] + 2 <= b->prof_ev[kind].size())
        hipEventRecord(b->prof_ev[kind][b->prof_used[kind]], s);
}
static void prof_end(te_batch* b, int kind, hipStream_t s) {
    if (((b->prof_kinds >> kind) & 1) && b->prof_used[kind] + 2 <= b->prof_ev[kind].size()) {
        hipEventRecord(b->prof_ev[kind][b->prof_used[kind] + 1], s);
        b->prof_used[kind] += 2;
    }
}
// After the stream is synchronised: sums the recorded pairs.
static void prof_collect(te_batch* b) {
    for (int k = 0; k < 4; k++) {
        b->prof_ms[k] = 0.0;
        b->prof_n[k] = b->prof_used[k] / 2;
        b->prof_each[k].clear();
        for (size_t i = 0; i + 1 < b->prof_used[k]; i += 2) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, b->prof_ev[k][i], b->prof_ev[k][i + 1]);
            b->prof_ms[k] += ms;
            b->prof_each[k].push_back(ms);
        }
    }
}

// The plan of the next round (or of the initial combine, or of a tuner's candidate shape).
// TE_BATCH_PLAN sets the planned units per CU to aim for (0: live units are never split; read
// at every plan, so a captured round keeps the value of its capture).
static void launch_plan(const te_batch* b, hipStream_t s) {
    int per_cu = kPlanUnitsPerCU;
    if (const char* e = std::getenv("TE_BATCH_PLAN")) per_cu = std::min(std::max(0, std::atoi(e)), 1 << 20);
    hipLaunchKernelGGL(batch_plan_kernel, dim3(1), dim3(kPlanThreads), 0, s, b->dev, b->ctx->num_cus * per_cu,
                       b->sell ? 1 : 0);
}

// The apply of iteration j for every instance: single-instance units, then shared-matrix units.
static void launch_batch_apply(const te_batch* b, hipStream_t s, int j) {
    const BatchDev& d = b->dev;
    if (!b->sell) {
        hipLaunchKernelGGL(batch_apply_kernel, dim3(b->grid_s), dim3(kBlock), 0, s, d, j);
        return;
    }
    if (d.n_sunits > 0) hipLaunchKernelGGL(batch_apply_sell_kernel, dim3(b->grid_s), dim3(kBlock), 0, s, d, j);
    if (d.n_gunits > 0) hipLaunchKernelGGL(batch_apply_group_kernel, dim3(b->grid_g), dim3(kBlock), 0, s, d, j);
}

// Enqueues one round on the context's stream (events only when timing is on, which also
// keeps the round out of a graph).
static void batch_round(te_batch* b, hipStream_t s) {
    const BatchDev& d = b->dev;
    launch_plan(b, s);
    for (int j = 0; j < b->m_max; j++) {
        prof_begin(b, 0, s);
        launch_batch_apply(b, s, j);
        hipLaunchKernelGGL(batch_finish_kernel<false>, dim3(b->n_inst), dim3(kBlock), 0, s, d, j);
        prof_end(b, 0, s);
        prof_begin(b, 1, s);
        hipLaunchKernelGGL(batch_project_kernel, dim3(b->grid_v), dim3(kBlock), 0, s, d, j);
        hipLaunchKernelGGL(batch_finish_kernel<true>, dim3(b->n_inst), dim3(kBlock), 0, s, d, j);
        prof_end(b, 1, s);
    }
    prof_begin(b, 2, s);
    launch_batch_step_control(b->consts, b->st, b->ks, b->bc, b->n_inst, b->n_samples, s);
    prof_end(b, 2, s);
}

static void batch_combine(te_batch* b, hipStream_t s) {
    const BatchDev& d = b->dev;
    prof_begin(b, 3, s);
    // one launch per group of kBatchObsPass observables (one pass over each stepped instance's
    // basis each); the last writes the new column 0 and the sampled states
    if (b->nobs == 0) hipLaunchKernelGGL(batch_combine_kernel<0>, dim3(b->grid_v), dim3(kBlock), 0, s, d, 0, 1);
    for (int o0 = 0; o0 < b->nobs; o0 += kBatchObsPass) {
        const int no = std::min(kBatchObsPass, b->nobs - o0), last = o0 + kBatchObsPass >= b->nobs ? 1 : 0;
        if (no == 1) hipLaunchKernelGGL(batch_combine_kernel<1>, dim3(b->grid_v), dim3(kBlock), 0, s, d, o0, last);
        else if (no == 2) hipLaunchKernelGGL(batch_combine_kernel<2>, dim3(b->grid_v), dim3(kBlock), 0, s, d, o0, last);
        else hipLaunchKernelGGL(batch_combine_kernel<4>, dim3(b->grid_v), dim3(kBlock), 0, s, d, o0, last);
    }
    hipLaunchKernelGGL(batch_obs_finish_kernel, dim3(b->n_inst), dim3(kBlock), 0, s, d);
    hipLaunchKernelGGL(batch_count_kernel, dim3(1), dim3(kBlock), 0, s, d);
    prof_end(b, 3, s);
}

// Drops the round graph (a failed capture or instantiation included): ends a capture still
// open on the stream, destroys what exists, clears the runtime's last error and leaves the
// batch launching kernel by kernel.
static void drop_graph(te_batch* b, hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(s, &g);
        if (g) hipGraphDestroy(g);
    }
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->use_graph = false;
    (void)hipGetLastError();
}

// Captures one round (Lanczos iterations, controller, combine) as a graph of kernel nodes
// and instantiates it. On any failure the batch falls back to direct launches (same results)
// and the run goes on; the failure is logged.
static void capture_round(te_batch* b, hipStream_t s) {
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        batch_round(b, s);
        batch_combine(b, s);
        e = hipGetLastError();
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(s, &g);
        if (e == hipSuccess) e = e2;
        b->graph = g;
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e == hipSuccess && !b->exec) e = hipErrorUnknown;
    }
    if (e != hipSuccess) {
        log_msg(2, "batched evolution: the round could not be captured as a HIP graph (%s); launching its "
                   "kernels one by one", hipGetErrorString(e));
        drop_graph(b, s);
    }
}

// Unit tables of a launch shape: SpMV units of unit_slices slices of one instance (sliced ELL;
// the CSR path keeps its fixed kBatchRB-row-block units), vector units of vec_rows rows (a
// multiple of kBatchChunk) of one instance.
static void shape_units(const te_batch* b, int unit_slices, int vec_rows, std::vector<int4>& su,
                        std::vector<int4>& gu, std::vector<int4>& vu) {
    su.clear();
    gu.clear();
    vu.clear();
    for (int i = 0; i < b->n_inst; i++) {
        const int g = b->group_of.empty() ? -1 : b->group_of[i];
        if (b->sell && g < 0)
            for (int sl = 0; sl < b->n_slices[i]; sl += unit_slices)
                su.push_back(make_int4(i, sl, std::min(sl + unit_slices, b->n_slices[i]), 0));
        if (b->sell && g >= 0 && b->groups[g].x == i)    // a group's units go with its first instance
            for (int sl = 0; sl < b->n_slices[i]; sl += unit_slices)
                gu.push_back(make_int4(i, sl, std::min(sl + unit_slices, b->n_slices[i]), g + 1));
        for (int r = 0; r < b->dim[i]; r += vec_rows) vu.push_back(make_int4(i, r, std::min(r + vec_rows, b->dim[i]), 0));
    }
}

static int shape_grid(const te_batch* b, int items, int blocks_per_cu) {
    const int g = std::min(items, b->ctx->num_cus * blocks_per_cu);
    return std::max(8, (g + 7) / 8 * 8);
}

// Uploads the unit tables of a shape (asynchronously on the context's stream) and points the
// launch arguments at them. Any captured round graph is dropped: it holds the old arguments.
static int apply_shape(te_batch* b, int unit_slices, int vec_rows, int spmv_bpc, int vec_bpc) {
    te_ctx* ctx = b->ctx;
    hipStream_t s = ctx->stream;
    std::vector<int4> su, gu, vu;
    shape_units(b, unit_slices, vec_rows, su, gu, vu);
    if (b->sell && !su.empty())
        TE_HIP(ctx, hipMemcpyAsync(b->sunit, su.data(), sizeof(int4) * su.size(), hipMemcpyHostToDevice, s));
    if (!gu.empty()) TE_HIP(ctx, hipMemcpyAsync(b->gunit, gu.data(), sizeof(int4) * gu.size(), hipMemcpyHostToDevice, s));
    TE_HIP(ctx, hipMemcpyAsync(b->vunit, vu.data(), sizeof(int4) * vu.size(), hipMemcpyHostToDevice, s));
    // the host vectors must outlive the copies
    TE_HIP(ctx, hipStreamSynchronize(s));
    if (b->sell) b->dev.n_sunits = (int)su.size();
    b->dev.n_gunits = (int)gu.size();
    b->dev.n_vunits = (int)vu.size();
    b->unit_slices = b->sell ? unit_slices : 0;
    b->vec_rows = vec_rows;
    b->spmv_bpc = spmv_bpc;
    b->vec_bpc = vec_bpc;
    b->grid_s = shape_grid(b, b->dev.n_sunits, spmv_bpc);
    b->grid_g = shape_grid(b, b->dev.n_gunits, spmv_bpc);
    b->grid_v = shape_grid(b, b->dev.n_vunits, vec_bpc);
    if (b->exec || b->graph) {
        if (b->exec) hipGraphExecDestroy(b->exec);
        if (b->graph) hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
    }
    return TE_OK;
}

struct BatchShape {
    int unit_slices, vec_rows, spmv_bpc, vec_bpc;
};

// Tuned shapes of this process, by instance-size key.
static std::mutex g_shape_mu;
static std::map<uint64_t, BatchShape> g_shape_cache;

// The tuner: times Lanczos iteration 0 of every live instance (the apply with its finish
// launch, then the projection with its finish launch) for a few candidate shapes and keeps the
// fastest of each kind (median of 3 timed launches after one untimed). Unit sizes and grids
// change no result bit (per-slice and per-chunk partials), so the choice only moves time. Runs
// with the run's initial state in place (U[:, 0] = psi0, scalars and flags uploaded); the
// caller uploads the Krylov scalars again afterwards (the timed iterations write them).
static int tune_shape(te_batch* b) {
    te_ctx* ctx = b->ctx;
    hipStream_t s = ctx->stream;
    const auto t0 = std::chrono::steady_clock::now();
    const int reps = 3;
    hipEvent_t ev[2 * (reps + 1)];
    for (auto& e : ev) TE_HIP(ctx, hipEventCreate(&e));
    auto time_kind = [&](bool spmv, float* med) -> int {
        launch_plan(b, s);                       // the candidate's units
        for (int r = 0; r <= reps; r++) {
            TE_HIP(ctx, hipEventRecord(ev[2 * r], s));
            if (spmv) {
                launch_batch_apply(b, s, 0);
                hipLaunchKernelGGL(batch_finish_kernel<false>, dim3(b->n_inst), dim3(kBlock), 0, s, b->dev, 0);
            } else {
                hipLaunchKernelGGL(batch_project_kernel, dim3(b->grid_v), dim3(kBlock), 0, s, b->dev, 0);
                hipLaunchKernelGGL(batch_finish_kernel<true>, dim3(b->n_inst), dim3(kBlock), 0, s, b->dev, 0);
            }
            TE_HIP(ctx, hipEventRecord(ev[2 * r + 1], s));
        }
        TE_HIP(ctx, hipGetLastError());
        TE_HIP(ctx, hipStreamSynchronize(s));
        float t[reps];
        for (int r = 1; r <= reps; r++) TE_HIP(ctx, hipEventElapsedTime(&t[r - 1], ev[2 * r], ev[2 * r + 1]));
        std::sort(t, t + reps);
        *med = t[reps / 2];
        return TE_OK;
    };
    int rc = TE_OK;
    BatchShape best{b->unit_slices, b->vec_rows, b->spmv_bpc, b->vec_bpc};
    // the apply needs its own w; the projection reads it: SpMV candidates first, then the
    // vector candidates on the w the last apply left
    if (b->sell) {
        float best_t = 1e30f;
        for (int us : {16, 32, 64, 128})
            for (int bpc : {4, 8, 16}) {
                if ((rc = apply_shape(b, us, best.vec_rows, bpc, best.vec_bpc)) != TE_OK) goto done;
                float t = 0.f;
                if ((rc = time_kind(true, &t)) != TE_OK) goto done;
                if (t < best_t) {
                    best_t = t;
                    best.unit_slices = us;
                    best.spmv_bpc = bpc;
                }
            }
    } else {
        float best_t = 1e30f;
        for (int bpc : {4, 8, 16}) {
            if ((rc = apply_shape(b, 0, best.vec_rows, bpc, best.vec_bpc)) != TE_OK) goto done;
            float t = 0.f;
            if ((rc = time_kind(true, &t)) != TE_OK) goto done;
            if (t < best_t) {
                best_t = t;
                best.spmv_bpc = bpc;
            }
        }
    }
    {
        float best_t = 1e30f;
        for (int vr : {1024, 2048, 4096, 8192})
            for (int bpc : {4, 8, 16}) {
                if ((rc = apply_shape(b, best.unit_slices, vr, best.spmv_bpc, bpc)) != TE_OK) goto done;
                float t = 0.f;          // (each projection moves w further; the data do not matter here)
                if ((rc = time_kind(false, &t)) != TE_OK) goto done;
                if (t < best_t) {
                    best_t = t;
                    best.vec_rows = vr;
                    best.vec_bpc = bpc;
                }
            }
    }
    rc = apply_shape(b, best.unit_slices, best.vec_rows, best.spmv_bpc, best.vec_bpc);
done:
    for (auto& e : ev) hipEventDestroy(e);
    b->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc == TE_OK) {
        b->shape_source = 1;
        std::lock_guard<std::mutex> lk(g_shape_mu);
        g_shape_cache[b->shape_key] = best;
    }
    return rc;
}

}  // namespace te

#undef TE_TRY
#define TE_TRY(expr)             \
    do {                         \
        int rc_ = (expr);        \
        if (rc_ != TE_OK) {      \
            te::batch_free(b);   \
            return rc_;          \
        }                        \
    } while (0)
#define TE_TRYH(call) TE_TRY((call) == hipSuccess ? TE_OK : te::fail(ctx, TE_ERR_HIP, "%s failed", #call))

extern "C" int te_batch_create(te_ctx* ctx, int n_inst, te_matrix* const* H, const double* const* psi0,
                               const te_evolve_params* params, const te_observable* obs, int n_obs,
                               te_batch** out) {
    using namespace te;
    if (!ctx || n_inst <= 0 || !H || !psi0 || !params || !out || n_obs < 0 || (n_obs > 0 && !obs))
        return TE_ERR_ARG;
    *out = nullptr;
    if (n_obs > kBatchMaxObs)
        return fail(ctx, TE_ERR_UNSUPPORTED, "batched evolution supports at most %d observables per instance",
                    kBatchMaxObs);
    if (!(params->t > 0) || !(params->sampling_step > 0)) return fail(ctx, TE_ERR_ARG, "t and samplingStep must be > 0");
    if (params->m < 1) return fail(ctx, TE_ERR_ARG, "Krylov dimension must be >= 1");
    hipSetDevice(ctx->device);
    te_batch* b = new te_batch();
    b->ctx = ctx;
    b->P = *params;
    b->n_inst = n_inst;
    b->nobs = n_obs;
    b->n_samples = (long long)std::floor(params->t / params->sampling_step) + 1;       // :57
    b->tolRate = params->tol / params->t;                                              // :287
    b->off.resize(n_inst);
    b->dim.resize(n_inst);
    b->m0.resize(n_inst);
    b->numErrEst.resize(n_inst);

    // sliced-ELL path when every instance has the copy (TE_BATCH_SELL=0: CSR row blocks)
    {
        const char* e = std::getenv("TE_BATCH_SELL");
        bool all = !(e && e[0] == '0');
        for (int i = 0; i < n_inst && all; i++)
            all = H[i] && H[i]->use_vd && H[i]->vd.sell && H[i]->vd.n_values <= kVdMaxValues;
        b->sell = all;
    }
    // default launch shape (the tuner may replace it on the first run; TE_BATCH_TUNE=0 keeps it).
    // Blocks per CU of the streaming kernels: 8 (0.247 s against 0.278 s at 4).
    b->spmv_bpc = b->vec_bpc = kBatchBlocksPerCU;
    b->vec_rows = kBatchVecRows;
    {
        const char* e = std::getenv("TE_BATCH_TUNE");
        b->tune = !(e && e[0] == '0');
    }
    b->perm.assign(n_inst, nullptr);
    b->iperm.assign(n_inst, nullptr);
    // instances: dimensions, Krylov dimensions, partial ranges
    std::vector<BatchInst> hinst(n_inst);
    std::vector<int4> hsu;
    int64_t total = 0;
    b->n_slices.assign(n_inst, 0);
    b->n_rblocks.assign(n_inst, 0);
    for (int i = 0; i < n_inst; i++) {
        te_matrix* A = H[i];
        if (!A || !psi0[i]) {
            batch_free(b);
            return fail(ctx, TE_ERR_ARG, "instance %d: null matrix or state", i);
        }
        if (A->ctx != ctx) {
            batch_free(b);
            return fail(ctx, TE_ERR_ARG, "instance %d: matrix belongs to another context", i);
        }
        if (A->n_rows == 0) {
            batch_free(b);
            return fail(ctx, TE_ERR_FATAL, "instance %d: Invalid Hilbertspace dimension", i);     // :39-43
        }
        if (A->n_cols != A->n_rows) {
            batch_free(b);
            return fail(ctx, TE_ERR_ARG, "instance %d: Hamiltonian must be square", i);
        }
        if (A->max_row_len + kStreamChunk - 1 > kStreamCap || A->n_row_blocks <= 0) {
            batch_free(b);
            return fail(ctx, TE_ERR_UNSUPPORTED, "instance %d: rows longer than %d entries", i,
                        kStreamCap - kStreamChunk + 1);
        }
        for (int o = 0; o < n_obs; o++) {
            const te_observable& ob = obs[(size_t)i * n_obs + o];
            if (ob.kind != TE_OBS_SPARSE || !ob.matrix || !ob.matrix->diagonal) {
                batch_free(b);
                return fail(ctx, TE_ERR_UNSUPPORTED, "instance %d: batched evolution takes diagonal sparse "
                                                     "observables only", i);
            }
            if (ob.matrix->n_rows != A->n_rows || ob.matrix->n_cols != A->n_cols) {
                batch_free(b);
                return fail(ctx, TE_ERR_FATAL, "instance %d: Incompatible dimensions", i);
            }
        }
        const int n = (int)A->n_rows;
        b->dim[i] = n;
        b->m0[i] = std::min(params->m, n);                                                   // :35
        if (b->m0[i] > kMaxKrylov) {
            batch_free(b);
            return fail(ctx, TE_ERR_UNSUPPORTED, "Krylov dimension %d exceeds %d", b->m0[i], kMaxKrylov);
        }
        b->m_max = std::max(b->m_max, b->m0[i]);
        double nrm1 = 0;
        TE_TRY(te_matrix_norm1(A, &nrm1));                                                   // :33
        b->numErrEst[i] = (double)n * nrm1 * DBL_EPSILON;                                    // :289
        b->off[i] = total;
        // 64-row alignment keeps every instance's columns on whole cache lines
        total += ((int64_t)n + 63) / 64 * 64;
        BatchInst& I = hinst[i];
        I.off = b->off[i];
        I.n = n;
        I.ptr = A->row_ptr;
        I.col = A->col;
        I.val = A->val;
        I.rb = A->row_block;
        I.su0 = (int)hsu.size();
        b->n_rblocks[i] = A->n_row_blocks;
        if (b->sell) {
            I.sell = sell_view_of(A);
            b->perm[i] = A->vd.perm;
            b->iperm[i] = A->vd.iperm;
            b->n_slices[i] = I.sell.n_slices;
            I.nsu = 0;
            I.sl0 = (int)b->total_groups;
            I.nsl = (I.sell.n_slices + kBatchGroup - 1) / kBatchGroup;
            b->total_slices += I.sell.n_slices;
            b->total_groups += I.nsl;
        } else {
            I.sell = SellView{};
            for (int r = 0; r < A->n_row_blocks; r += kBatchRB)
                hsu.push_back(make_int4(i, r, std::min(r + kBatchRB, A->n_row_blocks), 0));
            I.nsu = (int)hsu.size() - I.su0;
            I.sl0 = I.su0;
            I.nsl = I.nsu;
        }
        I.ch0 = (int)b->total_chunks;
        I.nch = (n + kBatchChunk - 1) / kBatchChunk;
        b->total_chunks += I.nch;
    }
    b->ld = total;
    // Instances that share a matrix (the same te_matrix) are served kBatchNV at a time by one pass
    // over its words (TE_BATCH_SHARED=0: every instance on its own); the results are the same bits.
    b->group_of.assign(n_inst, -1);
    if (b->sell) {
        const char* sh = std::getenv("TE_BATCH_SHARED");
        if (!(sh && sh[0] == '0')) {
            std::map<const te_matrix*, std::vector<int>> by;
            for (int i = 0; i < n_inst; i++) by[H[i]].push_back(i);
            std::vector<std::vector<int>> sets;
            for (int i = 0; i < n_inst; i++) {
                auto it = by.find(H[i]);
                if (it == by.end()) continue;
                sets.push_back(it->second);
                by.erase(it);
            }
            const char* nv_env = std::getenv("TE_BATCH_NV");            // group size sweeps: 2 or 4
            const size_t nv_max = (nv_env && std::atoi(nv_env) == 2) ? 2 : kBatchNV;
            for (const std::vector<int>& v : sets) {
                size_t k = 0;
                while (v.size() - k >= 2) {
                    const size_t take = v.size() - k >= nv_max ? nv_max : 2;
                    int4 g = make_int4(v[k], v[k + 1], take > 2 ? v[k + 2] : -1, take > 3 ? v[k + 3] : -1);
                    for (size_t q = 0; q < take; q++) b->group_of[v[k + q]] = (int)b->groups.size();
                    b->groups.push_back(g);
                    k += take;
                }
            }
        }
    }
    if (b->sell) {
        int64_t per = b->total_slices / std::max(8, ctx->num_cus * b->spmv_bpc);
        b->unit_slices = (int)std::max<int64_t>(kBatchSlicesMin, std::min<int64_t>(kBatchSlicesMax, per)) /
                         kBatchGroup * kBatchGroup;
    } else {
        b->unit_slices = 0;                 // CSR units are kBatchRB row blocks
    }
    if (b->total_chunks >= (1LL << 31) || (int64_t)hsu.size() >= (1LL << 31)) {
        batch_free(b);
        return fail(ctx, TE_ERR_UNSUPPORTED, "batched evolution: too many rows in all");
    }
    // instance-size key of the tuner's cache: the sorted (dimension, nnz) pairs, the storage and
    // the device
    {
        std::vector<std::pair<int64_t, int64_t>> sizes(n_inst);
        for (int i = 0; i < n_inst; i++) sizes[i] = {b->dim[i], H[i]->nnz};
        std::sort(sizes.begin(), sizes.end());
        uint64_t h = 1469598103934665603ull;
        auto mix = [&](uint64_t v) {
            for (int q = 0; q < 8; q++) {
                h ^= (v >> (8 * q)) & 0xffu;
                h *= 1099511628211ull;
            }
        };
        for (auto& p : sizes) {
            mix((uint64_t)p.first);
            mix((uint64_t)p.second);
        }
        mix(b->sell ? 1u : 0u);
        mix((uint64_t)b->groups.size());
        mix((uint64_t)ctx->device);
        mix((uint64_t)ctx->num_cus);
        b->shape_key = h;
    }
    const size_t ns = (size_t)b->n_samples;
    TE_TRY(balloc(ctx, &b->inst, n_inst));
    // unit tables sized for the smallest units: one group of slices (or kBatchRB row blocks) and
    // one chunk
    TE_TRY(balloc(ctx, &b->sunit, b->sell ? (size_t)b->total_groups : hsu.size()));
    TE_TRY(balloc(ctx, &b->vunit, (size_t)b->total_chunks));
    TE_TRY(balloc(ctx, &b->vgroup, std::max<size_t>(b->groups.size(), 1)));
    TE_TRY(balloc(ctx, &b->gunit, b->sell ? (size_t)b->total_groups : 1));
    TE_TRY(balloc(ctx, &b->U, (size_t)total * (b->m_max + 1)));
    TE_TRY(balloc(ctx, &b->psi0, (size_t)total));
    TE_TRY(balloc(ctx, &b->sampled, (size_t)total));
    TE_TRY(balloc(ctx, &b->tmp, (size_t)total));
    TE_TRY(balloc(ctx, &b->wcoef, (size_t)n_inst * kMaxKrylov));
    TE_TRY(balloc(ctx, &b->Y, (size_t)n_inst * ns * kMaxKrylov));
    TE_TRY(balloc(ctx, &b->spart, b->sell ? (size_t)b->total_groups : hsu.size()));
    TE_TRY(balloc(ctx, &b->vpart, (size_t)b->total_chunks));
    TE_TRY(balloc(ctx, &b->ks, n_inst));
    TE_TRY(balloc(ctx, &b->st, n_inst));
    TE_TRY(balloc(ctx, &b->bc, n_inst));
    TE_TRY(balloc(ctx, &b->consts, n_inst));
    TE_TRY(balloc(ctx, &b->diag, (size_t)std::max(n_obs, 1) * total));
    TE_TRY(balloc(ctx, &b->opart, (size_t)b->total_chunks * std::max(n_obs, 1) * ns));
    TE_TRY(balloc(ctx, &b->obs, (size_t)n_inst * std::max(n_obs, 1) * ns));
    TE_TRY(balloc(ctx, &b->nrm, n_inst));
    TE_TRY(balloc(ctx, &b->trace, (size_t)n_inst * kBatchTraceCap));
    TE_TRY(balloc(ctx, &b->n_live, 1));
    // a round's plan: at most one piece per group of slices and per chunk of rows
    TE_TRY(balloc(ctx, &b->lsunit, b->sell ? (size_t)b->total_groups : 1));
    TE_TRY(balloc(ctx, &b->lgunit, b->sell ? (size_t)b->total_groups : 1));
    TE_TRY(balloc(ctx, &b->lvunit, (size_t)b->total_chunks));
    TE_TRY(balloc(ctx, &b->n_plan, 3));
    TE_TRYH(hipMemsetAsync(b->n_plan, 0, sizeof(int) * 3, ctx->stream));
    TE_TRYH(hipHostMalloc((void**)&b->n_live_host, sizeof(int)));
    TE_TRY(upload_quadrature(ctx, b->quad));

    hipStream_t s = ctx->stream;
    TE_TRYH(hipMemcpyAsync(b->inst, hinst.data(), sizeof(BatchInst) * n_inst, hipMemcpyHostToDevice, s));
    if (!b->sell)
        TE_TRYH(hipMemcpyAsync(b->sunit, hsu.data(), sizeof(int4) * hsu.size(), hipMemcpyHostToDevice, s));
    if (!b->groups.empty())
        TE_TRYH(hipMemcpyAsync(b->vgroup, b->groups.data(), sizeof(int4) * b->groups.size(), hipMemcpyHostToDevice,
                               s));
    TE_TRYH(hipMemsetAsync(b->psi0, 0, sizeof(double2) * total, s));
    TE_TRYH(hipMemsetAsync(b->diag, 0, sizeof(double) * std::max(n_obs, 1) * total, s));
    std::vector<StepConsts> hc(n_inst);
    for (int i = 0; i < n_inst; i++) {
        // states and observables enter in the instance's own row order
        if (b->perm[i]) {
            TE_TRYH(hipMemcpyAsync(b->tmp + b->off[i], psi0[i], sizeof(double2) * b->dim[i], hipMemcpyHostToDevice, s));
            launch_gather(b->tmp + b->off[i], b->perm[i], b->psi0 + b->off[i], b->dim[i], s);
        } else {
            TE_TRYH(hipMemcpyAsync(b->psi0 + b->off[i], psi0[i], sizeof(double2) * b->dim[i], hipMemcpyHostToDevice,
                                   s));
        }
        for (int o = 0; o < n_obs; o++) {
            const double* dre = obs[(size_t)i * n_obs + o].matrix->diag_re;
            double* dst = b->diag + (size_t)o * total + b->off[i];
            if (b->perm[i]) launch_gather_re(dre, b->perm[i], dst, b->dim[i], s);
            else TE_TRYH(hipMemcpyAsync(dst, dre, sizeof(double) * b->dim[i], hipMemcpyDeviceToDevice, s));
        }
        StepConsts C = step_consts(b->P, b->tolRate, b->numErrEst[i], b->n_samples, b->quad);
        C.Y = b->Y + (size_t)i * ns * kMaxKrylov;
        C.wcoef = b->wcoef + (size_t)i * kMaxKrylov;
        C.trace = b->trace + (size_t)i * kBatchTraceCap;
        C.trace_cap = kBatchTraceCap;
        hc[i] = C;
    }
    TE_TRYH(hipMemcpyAsync(b->consts, hc.data(), sizeof(StepConsts) * n_inst, hipMemcpyHostToDevice, s));
    TE_TRYH(hipStreamSynchronize(s));
    // |‖psi0‖ - 1| <= tol (:45-48), per instance
    b->dev.inst = b->inst;
    b->dev.sampled = b->psi0;
    b->dev.nrm = b->nrm;
    hipLaunchKernelGGL(batch_norm_kernel, dim3(n_inst), dim3(kBlock), 0, s, b->dev);
    std::vector<double> n2(n_inst);
    TE_TRYH(hipMemcpyAsync(n2.data(), b->nrm, sizeof(double) * n_inst, hipMemcpyDeviceToHost, s));
    TE_TRYH(hipStreamSynchronize(s));
    for (int i = 0; i < n_inst; i++)
        if (std::fabs(std::sqrt(n2[i]) - 1.0) > params->tol) {
            batch_free(b);
            return fail(ctx, TE_ERR_FATAL, "instance %d: Initial vector is not normalized", i);
        }

    BatchDev& d = b->dev;
    d.inst = b->inst;
    d.sunit = b->sunit;
    d.vgroup = b->vgroup;
    d.gunit = b->gunit;
    d.n_sunits = (int)hsu.size();
    d.vunit = b->vunit;
    d.lsunit = b->lsunit;
    d.lgunit = b->lgunit;
    d.lvunit = b->lvunit;
    d.n_plan = b->n_plan;
    d.U = b->U;
    d.ld = total;
    d.alpha = make_double2(0.0, -1.0 * params->exp_factor);                               // :31
    d.tolRate = b->tolRate;
    d.ks = b->ks;
    d.st = b->st;
    d.bc = b->bc;
    d.spart = b->spart;
    d.vpart = b->vpart;
    d.wcoef = b->wcoef;
    d.Y = b->Y;
    d.n_samples = b->n_samples;
    d.diag = b->diag;
    d.nobs = n_obs;
    d.opart = b->opart;
    d.obs = b->obs;
    d.sampled = b->sampled;
    d.nrm = b->nrm;
    d.n_inst = n_inst;
    d.n_live = b->n_live;
    TE_TRY(apply_shape(b, b->unit_slices, b->vec_rows, b->spmv_bpc, b->vec_bpc));
    TE_TRYH(hipStreamSynchronize(s));
    *out = b;
    return TE_OK;
}

extern "C" void te_batch_destroy(te_batch* b) { te::batch_free(b); }

extern "C" int te_batch_info(const te_batch* b, size_t* packed_rows, size_t* n_samples, long long* rounds) {
    if (!b) return TE_ERR_ARG;
    if (packed_rows) *packed_rows = (size_t)b->ld;
    if (n_samples) *n_samples = (size_t)b->n_samples;
    if (rounds) *rounds = b->rounds;
    return TE_OK;
}

extern "C" int te_batch_set_graph(te_batch* b, int enable) {
    if (!b) return TE_ERR_ARG;
    b->use_graph = enable != 0;
    return TE_OK;
}

extern "C" int te_batch_set_shape(te_batch* b, int unit_slices, int vec_rows, int spmv_blocks_per_cu,
                                  int vec_blocks_per_cu) {
    using namespace te;
    if (!b) return TE_ERR_ARG;
    const int all = unit_slices | vec_rows | spmv_blocks_per_cu | vec_blocks_per_cu;
    if (all == 0) {                      // back to the tuner
        b->tune = true;
        return TE_OK;
    }
    if (unit_slices < kBatchGroup || unit_slices % kBatchGroup != 0 || vec_rows < kBatchChunk ||
        vec_rows % kBatchChunk != 0 || spmv_blocks_per_cu < 1 || vec_blocks_per_cu < 1 || spmv_blocks_per_cu > 64 ||
        vec_blocks_per_cu > 64)
        return fail(b->ctx, TE_ERR_ARG, "batched launch shape: unit_slices a positive multiple of %d, vec_rows a "
                                        "positive multiple of %d, blocks per CU in [1, 64]", kBatchGroup, kBatchChunk);
    hipSetDevice(b->ctx->device);
    const int rc = apply_shape(b, unit_slices, vec_rows, spmv_blocks_per_cu, vec_blocks_per_cu);
    if (rc != TE_OK) return rc;
    b->tune = false;
    b->shape_source = 3;
    return TE_OK;
}

extern "C" int te_batch_shape(const te_batch* b, int* unit_slices, int* vec_rows, int* spmv_blocks_per_cu,
                              int* vec_blocks_per_cu, int* source, double* tune_ms) {
    if (!b) return TE_ERR_ARG;
    if (unit_slices) *unit_slices = b->unit_slices;
    if (vec_rows) *vec_rows = b->vec_rows;
    if (spmv_blocks_per_cu) *spmv_blocks_per_cu = b->spmv_bpc;
    if (vec_blocks_per_cu) *vec_blocks_per_cu = b->vec_bpc;
    if (source) *source = b->shape_source;
    if (tune_ms) *tune_ms = b->tune_ms;
    return TE_OK;
}

extern "C" int te_batch_set_profiling(te_batch* b, int kinds) {
    if (!b || kinds < 0 || kinds > 15) return TE_ERR_ARG;
    b->prof_kinds = kinds;
    if (kinds && b->prof_ev[0].empty()) {
        hipSetDevice(b->ctx->device);
        for (auto& v : b->prof_ev) {
            v.resize(2 * te::kBatchProfPairs);
            for (hipEvent_t& e : v) TE_HIP(b->ctx, hipEventCreate(&e));
        }
    }
    return TE_OK;
}

extern "C" int te_batch_kernel_time(const te_batch* b, int kind, double* total_ms, size_t* launches) {
    if (!b || kind < 0 || kind > 3 || !total_ms || !launches) return TE_ERR_ARG;
    *total_ms = b->prof_ms[kind];
    *launches = b->prof_n[kind];
    return TE_OK;
}

extern "C" int te_batch_kernel_times(const te_batch* b, int kind, double* ms, size_t cap, size_t* launches) {
    if (!b || kind < 0 || kind > 3 || !launches || (cap > 0 && !ms)) return TE_ERR_ARG;
    const std::vector<double>& v = b->prof_each[kind];
    *launches = v.size();
    for (size_t i = 0; i < std::min(cap, v.size()); i++) ms[i] = v[i];
    return TE_OK;
}

#undef TE_TRY
#undef TE_TRYH

extern "C" int te_batch_run(te_batch* b, double* obs_values, double* const* states, te_result* results) {
    using namespace te;
    if (!b || !results) return TE_ERR_ARG;
    te_ctx* ctx = b->ctx;
    hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const int n_inst = b->n_inst;
    const size_t ns = (size_t)b->n_samples;

    for (int k = 0; k < 4; k++) b->prof_used[k] = 0;
    // initial state: U[:,0] = psi0, the sample at t = 0 (:305-313) through the combine pass
    std::vector<StepState> hst(n_inst);
    std::vector<BatchCtl> hbc(n_inst);
    std::vector<KrylovScalars> hks(n_inst);
    for (int i = 0; i < n_inst; i++) {
        StepState init{};
        init.t_step = 1.0 / 10.0;                                                      // :285
        init.m = b->m0[i];
        init.index_samples = 1;
        init.S = 1;
        init.sample_base = 0;
        init.write_w = 0;
        hst[i] = init;
        hbc[i].live = b->n_samples > 1 ? 1 : 0;
        hbc[i].stepped = 1;
        hbc[i].cm = 1;
        hbc[i].stall = 0;
        KrylovScalars k{};
        k.s[0] = 1.0;
        hks[i] = k;
    }
    std::vector<double2> e1((size_t)n_inst * kMaxKrylov, make_double2(0.0, 0.0));
    for (int i = 0; i < n_inst; i++) e1[(size_t)i * kMaxKrylov] = make_double2(1.0, 0.0);
    TE_HIP(ctx, hipMemcpyAsync(b->st, hst.data(), sizeof(StepState) * n_inst, hipMemcpyHostToDevice, s));
    TE_HIP(ctx, hipMemcpyAsync(b->bc, hbc.data(), sizeof(BatchCtl) * n_inst, hipMemcpyHostToDevice, s));
    TE_HIP(ctx, hipMemcpyAsync(b->ks, hks.data(), sizeof(KrylovScalars) * n_inst, hipMemcpyHostToDevice, s));
    for (int i = 0; i < n_inst; i++)
        TE_HIP(ctx, hipMemcpyAsync(b->Y + (size_t)i * ns * kMaxKrylov, e1.data() + (size_t)i * kMaxKrylov,
                                   sizeof(double2) * kMaxKrylov, hipMemcpyHostToDevice, s));
    TE_HIP(ctx, hipMemcpyAsync(b->U, b->psi0, sizeof(double2) * b->ld, hipMemcpyDeviceToDevice, s));
    if (b->tune) {
        // launch shape: the tuner's cached choice for these instance sizes, else measured now
        bool cached = false;
        BatchShape sh{};
        {
            std::lock_guard<std::mutex> lk(g_shape_mu);
            auto it = g_shape_cache.find(b->shape_key);
            if (it != g_shape_cache.end()) {
                cached = true;
                sh = it->second;
            }
        }
        int rc = cached ? apply_shape(b, sh.unit_slices, sh.vec_rows, sh.spmv_bpc, sh.vec_bpc) : tune_shape(b);
        if (rc != TE_OK) return rc;
        if (cached) {
            b->shape_source = 2;
            b->tune_ms = 0.0;
        }
        b->tune = false;
        // the tuner's iterations wrote the Krylov scalars: back to the run's start
        TE_HIP(ctx, hipMemcpyAsync(b->ks, hks.data(), sizeof(KrylovScalars) * n_inst, hipMemcpyHostToDevice, s));
        TE_HIP(ctx, hipStreamSynchronize(s));
    }
    TE_HIP(ctx, hipMemsetAsync(b->obs, 0, sizeof(double) * n_inst * std::max(b->nobs, 1) * ns, s));
    launch_plan(b, s);                           // every instance takes the t = 0 sample
    batch_combine(b, s);
    TE_HIP(ctx, hipGetLastError());
    if (b->P.fast_integration)
        log_msg(2, "Please note that a less accurate method for evaluating the error integral was used. We "
                   "recommend recomputing with accurate integration to establish the validity of the result.");

    // rounds until no instance is live
    const bool replay = b->use_graph && !b->prof_kinds;
    if (replay && !b->exec) capture_round(b, s);
    b->rounds = 0;
    TE_HIP(ctx, hipMemcpyAsync(b->n_live_host, b->n_live, sizeof(int), hipMemcpyDeviceToHost, s));
    TE_HIP(ctx, hipStreamSynchronize(s));
    const long long max_rounds = 1LL << 24;
    while (*b->n_live_host > 0) {
        if (b->rounds >= max_rounds) return fail(ctx, TE_ERR_FATAL, "batched evolution makes no progress");
        if (replay && b->exec) {
            TE_HIP(ctx, hipGraphLaunch(b->exec, s));
        } else {
            batch_round(b, s);
            batch_combine(b, s);
        }
        TE_HIP(ctx, hipGetLastError());
        TE_HIP(ctx, hipMemcpyAsync(b->n_live_host, b->n_live, sizeof(int), hipMemcpyDeviceToHost, s));
        TE_HIP(ctx, hipStreamSynchronize(s));
        b->rounds++;
    }

    // results (:445-504, :662-676)
    hipLaunchKernelGGL(batch_norm_kernel, dim3(n_inst), dim3(kBlock), 0, s, b->dev);
    std::vector<double> n2(n_inst);
    TE_HIP(ctx, hipMemcpyAsync(hst.data(), b->st, sizeof(StepState) * n_inst, hipMemcpyDeviceToHost, s));
    TE_HIP(ctx, hipMemcpyAsync(n2.data(), b->nrm, sizeof(double) * n_inst, hipMemcpyDeviceToHost, s));
    if (obs_values && b->nobs > 0)
        TE_HIP(ctx, hipMemcpyAsync(obs_values, b->obs, sizeof(double) * n_inst * b->nobs * ns, hipMemcpyDeviceToHost,
                                   s));
    if (states)
        for (int i = 0; i < n_inst; i++)
            if (states[i]) {
                const double2* src = b->sampled + b->off[i];
                if (b->iperm[i]) {                   // back to the basis order
                    launch_gather(src, b->iperm[i], b->tmp + b->off[i], b->dim[i], s);
                    src = b->tmp + b->off[i];
                }
                TE_HIP(ctx, hipMemcpyAsync(states[i], src, sizeof(double2) * b->dim[i], hipMemcpyDeviceToHost, s));
            }
    TE_HIP(ctx, hipStreamSynchronize(s));
    if (b->prof_kinds) prof_collect(b);
    int first_fatal = -1, fatal_code = 0;
    for (int i = 0; i < n_inst; i++) {
        const StepState& h = hst[i];
        te_result R{};
        R.dim = (size_t)b->dim[i];
        R.err = h.err;
        R.evolved_time = h.t_now;
        R.num_samples = (size_t)h.index_samples;
        R.n_steps = (size_t)h.n_steps;
        R.krylov_dim = (size_t)h.m;
        if (h.fatal) {
            R.status_code = -1;
            if (first_fatal < 0) {
                first_fatal = i;
                fatal_code = h.fatal;
            }
        } else {
            R.status_code = final_status(h, b->numErrEst[i], b->P.tol, std::sqrt(n2[i]));
        }
        results[i] = R;
    }
    if (first_fatal >= 0) {
        const char* why = fatal_code == 1 ? "No small enough time step found to meet tolerance requirements."
                          : fatal_code == 2 ? "eigen-solver did not converge"
                                            : "time evolution makes no progress (t_step = 0)";
        return fail(ctx, TE_ERR_FATAL, "instance %d: %s", first_fatal, why);
    }
    return TE_OK;
}
