// Adjoint gradient of L = <psi(T)|O|psi(T)> with respect to the coefficients theta_p of a
// Hamiltonian H = H_0 + sum_p theta_p G_p (no counterpart in the reference, which only
// evolves forward: core/krylovTimeEvolver.cpp:261-512). With psi(s) = exp(-i e H s) psi0
// (e = expFactor) and lambda(s) = exp(i e H (T - s)) O psi(T),
//   dL/dtheta_p = 2 e int_0^T Im <lambda(s)| G_p |psi(s)> ds.
//
// Forward: the plain method (same Lanczos kernels, step controller and error bound) with the
// step schedule, every step's Lanczos scalars and either every Krylov basis (checkpoint
// interval K = 0) or the boundary states psi(t_k) for k % K == 0 (K >= 1) kept in HBM.
// Backward, step by step from T on the forward's own schedule:
//  * lambda is back-evolved: a Lanczos space of +i e H from lambda(t_{k+1});
//  * psi on the step comes from the kept forward basis (K = 0), or is back-evolved alongside
//    lambda from psi(t_{k+1}) and replaced by the forward checkpoint every K steps, which
//    bounds its drift (K >= 1);
//  * both spaces are decomposed in one kernel, which picks a Gauss-Legendre rule for the
//    step from their Ritz spectra (adjoint_coef_kernel);
//  * the node states of both are formed in one pass over each basis, stored node-interleaved
//    (a row holds 8 consecutive nodes, 128 bytes), and one labelled term-wise pass per group
//    of 8 nodes evaluates every off-diagonal generator at once; diagonal generators reduce to
//    one running vector Z(r) = sum_nodes w Im(conj(lambda(r)) psi(r)), dotted with each
//    generator's diagonal at the end.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "te_internal.h"
#include "te_krylov.h"
#include "te_reduce.h"

namespace te {

constexpr int kNodeGroup = 8;       // nodes per node-interleaved block
constexpr int kMatNodes = 64;       // nodes per materialisation pass (8 groups)
constexpr int kAdjMaxLabels = 64;   // generators (LDS accumulators of the term-wise pass)
constexpr int kAdjMaxSteps = 1 << 16;
constexpr double kAdjRankTol = 1e-14;   // relative Frobenius residual of a step's rank form

__global__ void diag_mul_kernel(const double* __restrict__ d, const double2* __restrict__ x, double2* __restrict__ y,
                                int64_t n) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        y[r] = make_double2(d[r] * x[r].x, d[r] * x[r].y);
}

// After |v|^2 landed in res[0].x: the Lanczos start v_0 = v / |v| (s_0 = 1 / |v|), scale = |v|.
__global__ void start_norm_kernel(const double2* res, KrylovScalars* ks, double* scale) {
    const double nv = sqrt(res[0].x);
    ks->s[0] = nv > 0.0 ? 1.0 / nv : 1.0;
    ks->broken = 0;
    *scale = nv;
}

// Node states of one basis for the pass of nodes [node0, node0 + 64) of the step (the node count
// and the basis' Krylov dimension are read from the step record, written by the coefficient
// kernels on the device): out[g][r][q'] = sum_j U[j][r] coef[node0 + 8 g + q'][j] (zero past the
// step's node count; out holds the pass's blocks); end[r] = sum_j U[j][r] cend[j] when end is
// given. A wavefront forms one 8-node block for a tile of 256 rows, kMatRows rows a lane: every
// coefficient is an LDS broadcast (one address for the whole wavefront) that feeds 4 kMatRows
// FP64 FMAs, and the basis rows load two columns ahead. Work items are (tile, block) pairs, the
// blocks of one tile on neighbouring wavefronts, which read the same basis rows together (the
// pass re-reads the basis once per block, from L2: splitting a block over two wavefronts doubled
// that and was slower). A lane writes one 128-byte block per row.
struct MatArgs {
    const double2* U;
    int64_t ld, n;
    int basis;                // 0: psi, 1: lambda (its Krylov dimension: step->m[basis])
    const double2* coef;      // kAdjMaxNodes x kMaxKrylov
    const double2* cend;      // kMaxKrylov (with end)
    const AdjStep* step;
    int node0;
    double2* out;             // node blocks, n x 8 each
    double2* end;
};

constexpr int kMatRows = 4;        // rows formed by one lane
constexpr int kMatTile = 64 * kMatRows;

template <int MS>                  // the largest Krylov dimension of the launch (LDS table rows)
__global__ __launch_bounds__(256) void materialise_nodes_kernel(MatArgs a) {
    __shared__ double2 cs[MS * (kMatNodes + 1)];              // cs[j * (nq + 1) + q], q = nq: cend
    const int nn = a.step->n_nodes;
    const int m = a.step->m[a.basis];
    const int nloc = max(0, min(kMatNodes, nn - a.node0));
    const int ng = (nloc + kNodeGroup - 1) / kNodeGroup;     // blocks written by this pass
    const int nq = ng * kNodeGroup;
    const int ngi = max(ng, 1);                               // blocks per tile in the work items
    for (int i = threadIdx.x; i < m * (nq + 1); i += blockDim.x) {
        const int j = i / (nq + 1), q = i % (nq + 1);
        double2 c = make_double2(0.0, 0.0);
        if (q < nloc) c = a.coef[(size_t)(a.node0 + q) * kMaxKrylov + j];
        else if (q == nq && a.end) c = a.cend[j];
        cs[i] = c;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int waves = blockDim.x >> 6;
    const int64_t items = (a.n + kMatTile - 1) / kMatTile * ngi;   // (tile, block) pairs
    for (int64_t it = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * waves + (threadIdx.x >> 6))); it < items;
         it += (int64_t)gridDim.x * waves) {
        const int64_t tile = it / ngi;
        const int g = (int)(it % ngi);
        const bool with_end = a.end && g == 0;
        if (g >= ng && !with_end) continue;
        int r[kMatRows];                                   // rows < 2^31 (te_adjoint_create)
#pragma unroll
        for (int i = 0; i < kMatRows; i++) r[i] = (int)(tile * kMatTile) + i * 64 + lane;
        double2 acc[kMatRows][kNodeGroup], ae[kMatRows];
#pragma unroll
        for (int i = 0; i < kMatRows; i++) {
            ae[i] = make_double2(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < kNodeGroup; q++) acc[i][q] = make_double2(0.0, 0.0);
        }
        auto load_u = [&](int j, double2* u) {
#pragma unroll
            for (int i = 0; i < kMatRows; i++)
                u[i] = (j < m && r[i] < a.n) ? a.U[(size_t)j * a.ld + r[i]] : make_double2(0.0, 0.0);
        };
        double2 u0[kMatRows], u1[kMatRows];
        load_u(0, u0);
        load_u(1, u1);
        const int gq = g < ng ? g * kNodeGroup : 0;
        for (int j = 0; j < m; j++) {
            double2 u2[kMatRows];
            load_u(j + 2, u2);
            const double2* cj = cs + j * (nq + 1) + gq;
#pragma unroll
            for (int q = 0; q < kNodeGroup; q++) {
                const double2 c = cj[q];
#pragma unroll
                for (int i = 0; i < kMatRows; i++) {
                    acc[i][q].x += u0[i].x * c.x - u0[i].y * c.y;
                    acc[i][q].y += u0[i].x * c.y + u0[i].y * c.x;
                }
            }
            if (with_end) {
                const double2 c = cs[j * (nq + 1) + nq];
#pragma unroll
                for (int i = 0; i < kMatRows; i++) {
                    ae[i].x += u0[i].x * c.x - u0[i].y * c.y;
                    ae[i].y += u0[i].x * c.y + u0[i].y * c.x;
                }
            }
#pragma unroll
            for (int i = 0; i < kMatRows; i++) {
                u0[i] = u1[i];
                u1[i] = u2[i];
            }
        }
#pragma unroll
        for (int i = 0; i < kMatRows; i++) {
            if (r[i] >= a.n) continue;
            if (g < ng) {
                double2* o = a.out + ((size_t)g * a.n + r[i]) * kNodeGroup;
#pragma unroll
                for (int q = 0; q < kNodeGroup; q++) o[q] = acc[i][q];
            }
            if (with_end) a.end[r[i]] = ae[i];
        }
    }
}

// The same pass on the matrix cores: per 16-row tile the complex product U[:, rows]^T C (m
// basis columns by the pass's nodes, and the end vector as one more node) as a real product of
// the 16 x 2m matrix A (A[r][2j] = Re U[j][r], A[r][2j+1] = Im U[j][r]) by the 2m x 2(nodes + 1)
// matrix B (B[2j][2q] = Re C[q][j], B[2j][2q+1] = Im C[q][j], B[2j+1][2q] = -Im C[q][j],
// B[2j+1][2q+1] = Re C[q][j]), in output tiles of 16 columns (8 nodes) by
// v_mfma_f64_16x16x4f64: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16] of each 4-deep
// step and D[4 i + l / 16][l % 16] of the tile. A wavefront loads the tile's 2m x 16 basis
// values once (lanes 0-31 one column, 32-63 the next: 256 contiguous bytes each) and reads B
// from the coefficient table in LDS. The sums run over j in a different order from the vector
// form's (4-deep steps): the node states agree with it to rounding.
typedef double d4v __attribute__((ext_vector_type(4)));

template <int MS>
__global__ __launch_bounds__(256) void materialise_mfma_kernel(MatArgs a) {
    __shared__ double2 cs[MS * (kMatNodes + 1)];              // cs[j * (nq + 1) + q], q = nq: cend
    const int nn = a.step->n_nodes;
    const int m = a.step->m[a.basis];
    const int nloc = max(0, min(kMatNodes, nn - a.node0));
    const int ng = (nloc + kNodeGroup - 1) / kNodeGroup;     // blocks written by this pass
    const int nq = ng * kNodeGroup;
    for (int i = threadIdx.x; i < m * (nq + 1); i += blockDim.x) {
        const int j = i / (nq + 1), q = i % (nq + 1);
        double2 c = make_double2(0.0, 0.0);
        if (q < nloc) c = a.coef[(size_t)(a.node0 + q) * kMaxKrylov + j];
        else if (q == nq && a.end) c = a.cend[j];
        cs[i] = c;
    }
    __syncthreads();
    const int ncol = 2 * (nq + (a.end ? 1 : 0));                // real output columns
    const int ntile = (ncol + 15) / 16;
    if (ntile == 0) return;
    const int ks = (m + 1) / 2;                                 // 4-deep steps (2 basis columns each)
    const int lane = threadIdx.x & 63;
    const int waves = blockDim.x >> 6;
    const int jl = lane >> 5, part = (lane >> 4) & 1;           // A and B: column 2s + jl, re/im part
    const int64_t tiles = (a.n + 15) / 16;
    const double* U = reinterpret_cast<const double*>(a.U);
    for (int64_t t = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * waves + (threadIdx.x >> 6))); t < tiles;
         t += (int64_t)gridDim.x * waves) {
        const int64_t row = t * 16 + (lane & 15);
        double av[MS / 2 + 1];
#pragma unroll
        for (int st = 0; st < MS / 2 + 1; st++) {
            const int j = 2 * st + jl;
            av[st] = (st < ks && j < m && row < a.n) ? U[((size_t)j * a.ld + row) * 2 + part] : 0.0;
        }
        for (int T = 0; T < ntile; T++) {
            const int col = 16 * T + (lane & 15);
            const int q = col >> 1, p = col & 1;
            const bool qin = q <= nq;                           // q == nq: the end vector
            d4v acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int st = 0; st < MS / 2 + 1; st++) {
                if (st < ks) {                                  // the same for the whole wavefront
                    const int j = 2 * st + jl;
                    double b = 0.0;
                    if (qin && j < m) {
                        const double2 c = cs[j * (nq + 1) + q];
                        b = part == 0 ? (p == 0 ? c.x : c.y) : (p == 0 ? -c.y : c.x);
                    }
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[st], b, acc, 0, 0, 0);
                }
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t r = t * 16 + 4 * i + (lane >> 4);
                if (r >= a.n || !qin) continue;
                if (q < nq) {
                    double* o = reinterpret_cast<double*>(a.out + ((size_t)(q >> 3) * a.n + r) * kNodeGroup + (q & 7));
                    o[p] = acc[i];
                } else if (a.end) {
                    reinterpret_cast<double*>(a.end + r)[p] = acc[i];
                }
            }
        }
    }
}

// The node blocks of one materialisation pass (the step's nodes [node0, node0 + 64), count from
// the step record): for every row r and node q, L_q = w_q conj(lambda_q(r)),
//   Z(r) += sum_q Im(L_q psi_q(r))                              (diagonal generators)
//   acc[label] += Im(t sum_q L_q psi_q(c))  per entry (r, c, label, t)   (off-diagonal ones)
// The labelled entries are held as sliced ELL (64-row slices, slot-major, padded to the slice's
// longest row, row lengths predicate the padding) with values through a table of the distinct
// values in LDS (or stored when there are too many); a wavefront takes a slice, lane = row, and
// goes through the pass's 8-node blocks GP at a time (GP blocks gathered per entry), loading the
// next slot's column, label and value index before the current slot's gathers. Per-thread label
// accumulators live in LDS; block sums per label go to partials, which termwise_finish_kernel
// adds (in block order) to out[label].
constexpr int kTermMaxValues = 1024;

struct TermArgs {
    int64_t n;
    int n_slices;               // 0: no off-diagonal entries
    const int* soff;            // n_slices + 1 slot offsets (multiples of 64)
    const int* scol;
    const unsigned char* slab;  // compact label (generators with off-diagonal entries)
    const unsigned short* svid; // value index (table) ...
    const double2* sval;        // ... or the value
    const unsigned short* slen; // entries per row
    const double2* vtab;
    int n_values;
    int n_labels;               // compact labels
    const double2* psi;         // the pass's node blocks, n x 8 each
    const double2* lam;
    const double* w;            // kAdjMaxNodes weights
    const AdjStep* step;
    int node0;
    double* zsum;
    double* partials;           // gridDim.x x n_labels
    double* out;                // n_labels running sums (termwise_finish_kernel adds to them)
};

template <int LCAP, bool TABLE, int GP>
__global__ __launch_bounds__(256) void termwise_kernel(TermArgs a) {
    __shared__ double acc[LCAP * 256];
    __shared__ double2 vt[TABLE ? kTermMaxValues : 1];
    __shared__ double wq[kMatNodes];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nloc = max(0, min(kMatNodes, a.step->n_nodes - a.node0));
    const int ng = (nloc + kNodeGroup - 1) / kNodeGroup;
    for (int l = 0; l < a.n_labels; l++) acc[l * 256 + tid] = 0.0;
    if (TABLE)
        for (int i = tid; i < a.n_values; i += 256) vt[i] = a.vtab[i];
    if (tid < kMatNodes) wq[tid] = tid < nloc ? a.w[a.node0 + tid] : 0.0;
    __syncthreads();
    const int n_sl = (int)((a.n + 63) / 64);
    const size_t gstride = (size_t)a.n * kNodeGroup;
    for (int sl = blockIdx.x * 4 + wv; sl < n_sl; sl += gridDim.x * 4) {
        const int64_t r = (int64_t)sl * 64 + lane;
        const int len = (r < a.n && a.n_slices) ? a.slen[r] : 0;
        const int o = a.n_slices ? a.soff[sl] : 0;
        const int ls = a.n_slices ? (a.soff[sl + 1] - o) >> 6 : 0;
        double z = 0.0;
        for (int g0 = 0; g0 < ng; g0 += GP) {
            double2 L[GP][kNodeGroup];
#pragma unroll
            for (int h = 0; h < GP; h++) {
                const int g = g0 + h;
#pragma unroll
                for (int q = 0; q < kNodeGroup; q++) L[h][q] = make_double2(0.0, 0.0);
                if (r < a.n && g < ng) {
                    const double2* lr = a.lam + g * gstride + (size_t)r * kNodeGroup;
                    const double2* pr = a.psi + g * gstride + (size_t)r * kNodeGroup;
#pragma unroll
                    for (int q = 0; q < kNodeGroup; q++) {
                        const double2 l = lr[q];
                        const double wv_ = wq[g * kNodeGroup + q];
                        L[h][q] = make_double2(wv_ * l.x, -wv_ * l.y);
                        const double2 p = pr[q];
                        z += L[h][q].x * p.y + L[h][q].y * p.x;
                    }
                }
            }
            if (ls == 0) continue;
            const bool two = GP > 1 && g0 + 1 < ng;
            int c = 0, lb = 0, vi = 0;
            double2 tv = make_double2(0.0, 0.0);
            {
                const size_t k = (size_t)o + lane;
                c = a.scol[k];
                lb = a.slab[k];
                if (TABLE) vi = a.svid[k];
                else tv = a.sval[k];
            }
            for (int j = 0; j < ls; j++) {
                const int cc = c, lbl = lb;
                const double2 t = TABLE ? vt[vi] : tv;
                if (j + 1 < ls) {                          // next slot's indices in flight
                    const size_t k = (size_t)o + (size_t)(j + 1) * 64 + lane;
                    c = a.scol[k];
                    lb = a.slab[k];
                    if (TABLE) vi = a.svid[k];
                    else tv = a.sval[k];
                }
                if (j < len) {
                    double sr = 0.0, si = 0.0;
#pragma unroll
                    for (int h = 0; h < GP; h++) {
                        if (h > 0 && !two) break;
                        const double2* pc = a.psi + (g0 + h) * gstride + (size_t)cc * kNodeGroup;
#pragma unroll
                        for (int q = 0; q < kNodeGroup; q++) {
                            const double2 p = pc[q];
                            sr += L[h][q].x * p.x - L[h][q].y * p.y;
                            si += L[h][q].x * p.y + L[h][q].y * p.x;
                        }
                    }
                    acc[lbl * 256 + tid] += t.x * si + t.y * sr;
                }
            }
        }
        if (r < a.n && ng > 0) a.zsum[r] += z;
    }
    __shared__ double2 sh[4];
    for (int l = 0; l < a.n_labels; l++) {
        const double2 s = block_sum2<256>(make_double2(acc[l * 256 + tid], 0.0), sh);
        if (tid == 0) a.partials[(size_t)blockIdx.x * a.n_labels + l] = s.x;
    }
}

// One block after termwise_kernel: out[l] += the nparts blocks' partials of label l, summed in
// block order (in a launch of its own: a last-block reduction inside the term-wise pass would
// cost each of its blocks an agent-scope release, an L2 write-back on this multi-XCD part).
__global__ __launch_bounds__(256) void termwise_finish_kernel(const double* __restrict__ partials, int nparts,
                                                              int n_labels, double* out) {
    __shared__ double2 sh[4];
    const int tid = threadIdx.x;
    for (int l = 0; l < n_labels; l++) {
        double t = 0.0;
        for (int b = tid; b < nparts; b += 256) t += partials[(size_t)b * n_labels + l];
        const double2 s = block_sum2<256>(make_double2(t, 0.0), sh);
        if (tid == 0) out[l] += s.x;
    }
}

// out[p] = sum_r diag[p][r] z[r] for every generator p (its diagonal part; zero rows where it
// has none): block partials per generator, the last block sums them in block order.
__global__ __launch_bounds__(256) void diag_grad_kernel(const double* __restrict__ diag, const double* __restrict__ z,
                                                        int64_t n, int P, double* partials, unsigned int* ticket,
                                                        double* out) {
    __shared__ double2 sh[4];
    const int tid = threadIdx.x;
    for (int p = 0; p < P; p++) {
        double acc = 0.0;
        for (int64_t r = (int64_t)blockIdx.x * 256 + tid; r < n; r += (int64_t)gridDim.x * 256)
            acc += diag[(size_t)p * n + r] * z[r];
        const double2 s = block_sum2<256>(make_double2(acc, 0.0), sh);
        if (tid == 0) partials[(size_t)blockIdx.x * P + p] = s.x;
    }
    if (last_block_arrive(ticket)) {
        for (int p = 0; p < P; p++) {
            double t = 0.0;
            for (int b = tid; b < (int)gridDim.x; b += 256) t += partials[(size_t)b * P + p];
            const double2 s = block_sum2<256>(make_double2(t, 0.0), sh);
            if (tid == 0) out[p] = s.x;
        }
        if (tid == 0) *ticket = 0;
    }
}

template <bool TABLE, int GP>
static void launch_termwise_t(const TermArgs& ta, int grid, hipStream_t s) {
    if (ta.n_labels <= 8) hipLaunchKernelGGL((termwise_kernel<8, TABLE, GP>), dim3(grid), dim3(256), 0, s, ta);
    else if (ta.n_labels <= 16) hipLaunchKernelGGL((termwise_kernel<16, TABLE, GP>), dim3(grid), dim3(256), 0, s, ta);
    else if (ta.n_labels <= 32) hipLaunchKernelGGL((termwise_kernel<32, TABLE, GP>), dim3(grid), dim3(256), 0, s, ta);
    else hipLaunchKernelGGL((termwise_kernel<64, TABLE, GP>), dim3(grid), dim3(256), 0, s, ta);
}

// One term-wise pass over a materialisation pass's node blocks (gp: blocks gathered per entry,
// 1 or 2), then its finish launch.
static void launch_termwise(const TermArgs& ta, int grid, int gp, hipStream_t s) {
    if (ta.vtab) {
        if (gp > 1) launch_termwise_t<true, 2>(ta, grid, s);
        else launch_termwise_t<true, 1>(ta, grid, s);
    } else {
        if (gp > 1) launch_termwise_t<false, 2>(ta, grid, s);
        else launch_termwise_t<false, 1>(ta, grid, s);
    }
    hipLaunchKernelGGL(termwise_finish_kernel, dim3(1), dim3(256), 0, s, ta.partials, grid, ta.n_labels, ta.out);
}

}  // namespace te

using namespace te;

struct te_adjoint {
    te_ctx* ctx = nullptr;
    te_matrix* H = nullptr;
    te_matrix* O = nullptr;
    int P = 0;
    te_evolve_params prm{};
    int K = 0;
    int64_t n = 0;
    int m0 = 0;
    double tolRate = 0, numErrEst = 0;
    const int* perm = nullptr;      // H's locality order (new -> old), null for the basis order
    const int* iperm = nullptr;
    int grid = 8;
    int pgrid = 8;
    // generators in the evolver's order
    int* soff = nullptr;            // labelled off-diagonal entries, sliced ELL
    int* scol = nullptr;
    unsigned char* slab = nullptr;
    unsigned short* svid = nullptr;
    double2* sval = nullptr;
    unsigned short* slen = nullptr;
    double2* vtab = nullptr;
    int n_values = 0, n_slices = 0;
    int n_off = 0;                  // generators with off-diagonal entries (compact labels)
    std::vector<int> off_label;     // generator -> compact label, or -1
    int64_t g_nnz = 0;
    double* gdiag = nullptr;        // P x n (Re of the diagonal part)
    double* odiag = nullptr;        // n
    // forward
    std::vector<double2*> bases;    // K = 0: one (m0 + 1) x n basis per step
    double2* U = nullptr;           // K >= 1
    std::vector<double2*> ckpt;     // K >= 1: psi(t_k), k % K == 0 (kept across runs, grown as needed)
    size_t n_ckpt = 0;              // checkpoints written by the last forward
    double2* psiT = nullptr;
    double2* scratch = nullptr;     // currentVec of the last step
    double2* psi0 = nullptr;
    KrylovScalars* ks = nullptr;
    KrylovScalars* ks_hist = nullptr;
    StepState* st = nullptr;
    StepState* st_host = nullptr;
    double2 *wcoef = nullptr, *Y = nullptr;
    double* trace = nullptr;
    double* obs = nullptr;          // 2 samples of <O>
    DevQuad quad;
    ReduceScratch rs;
    std::vector<double> t_steps;
    std::vector<int> m_steps;
    // backward
    double2 *Vl = nullptr, *Vp = nullptr;
    KrylovScalars *ksl = nullptr, *ksp = nullptr;
    double* scale = nullptr;        // [0] psi start norm, [1] lambda start norm
    double2* gl = nullptr;
    int* gl_off = nullptr;
    double2 *coef0 = nullptr, *coef1 = nullptr, *cend0 = nullptr, *cend1 = nullptr;
    double* w = nullptr;
    hipStream_t s2 = nullptr;       // term-wise passes beside the next step's coefficient kernels
    hipEvent_t ev_lanczos = nullptr, ev_term = nullptr;
    AdjStep* steps = nullptr;       // one record per step of the last backward
    int steps_cap = 0;
    AdjStep* step_host = nullptr;
    double2 *nodes_psi = nullptr, *nodes_lam = nullptr;
    double2 *nodes_psi1 = nullptr, *nodes_lam1 = nullptr;   // odd steps in low-rank form
    double2 *psi_next = nullptr, *lam_next = nullptr;
    double* zsum = nullptr;
    double* offacc = nullptr;       // P
    double* dgrad = nullptr;        // P
    double* partials = nullptr;
    unsigned int* ticket = nullptr;
    // last run
    te_adjoint_info info{};
};

namespace te {

static void adj_free_bases(te_adjoint* a) {
    for (double2* p : a->bases)
        if (p) hipFree(p);
    a->bases.clear();
    for (double2* p : a->ckpt)
        if (p) hipFree(p);
    a->ckpt.clear();
}

static void adj_free(te_adjoint* a) {
    if (!a) return;
    hipSetDevice(a->ctx->device);
    if (a->s2) hipStreamSynchronize(a->s2);        // no term-wise pass left reading the buffers
    adj_free_bases(a);
    void* ptrs[] = {a->soff, a->scol, a->slab, a->svid, a->sval, a->slen, a->vtab, a->gdiag,  a->odiag,  a->U,      a->psiT,
                    a->scratch, a->psi0, a->ks,   a->ks_hist, a->st,  a->wcoef, a->Y,      a->trace,
                    a->obs,  a->Vl,      a->Vp,   a->ksl,  a->ksp,    a->scale,  a->gl,     a->gl_off,
                    a->coef0, a->coef1,  a->cend0, a->cend1, a->w,    a->steps,  a->nodes_psi, a->nodes_lam,
                    a->nodes_psi1, a->nodes_lam1,
                    a->psi_next, a->lam_next, a->zsum, a->offacc, a->dgrad, a->partials, a->ticket};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (a->st_host) hipHostFree(a->st_host);
    if (a->step_host) hipHostFree(a->step_host);
    if (a->ev_lanczos) hipEventDestroy(a->ev_lanczos);
    if (a->ev_term) hipEventDestroy(a->ev_term);
    if (a->s2) hipStreamDestroy(a->s2);
    free_quadrature(a->quad);
    reduce_scratch_free(a->rs);
    delete a;
}

template <class T>
static int aalloc(te_ctx* ctx, T** p, size_t count) {
    TE_HIP(ctx, hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1)));
    return TE_OK;
}

// Gauss-Legendre orders 1..kGlMax on [-1, 1]: nodes ascending with their weights.
static void gl_table(std::vector<double2>& nodes, std::vector<int>& off) {
    off.assign(kGlMax + 2, 0);
    for (int n = 1; n <= kGlMax; n++) {
        off[n] = (int)nodes.size();
        std::vector<double2> row;
        for (int i = 1; i <= n; i++) {
            double z = std::cos(M_PI * (i - 0.25) / (n + 0.5));
            double dp = 1.0;
            for (int it = 0; it < 100; it++) {
                double p0 = 1.0, p1 = z;
                for (int k = 2; k <= n; k++) {
                    const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                dp = n * (z * p1 - p0) / (z * z - 1.0);
                const double dz = p1 / dp;
                z -= dz;
                if (std::fabs(dz) < 1e-16) break;
            }
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; k++) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            row.push_back(make_double2(z, 2.0 / ((1.0 - z * z) * dp * dp)));
        }
        std::sort(row.begin(), row.end(), [](const double2& x, const double2& y) { return x.x < y.x; });
        nodes.insert(nodes.end(), row.begin(), row.end());
    }
    off[kGlMax + 1] = (int)nodes.size();
}

}  // namespace te

#define TE_TRY(expr)           \
    do {                       \
        int rc_ = (expr);      \
        if (rc_ != TE_OK) {    \
            te::adj_free(a);   \
            return rc_;        \
        }                      \
    } while (0)
#define TE_TRYH(call) TE_TRY((call) == hipSuccess ? TE_OK : te::fail(ctx, TE_ERR_HIP, "%s failed", #call))

extern "C" int te_adjoint_create(te_ctx* ctx, te_matrix* H, int n_params, te_matrix* const* generators,
                                 te_matrix* observable, const te_evolve_params* params, int checkpoint_interval,
                                 te_adjoint** out) {
    using namespace te;
    if (!ctx || !H || n_params < 0 || (n_params > 0 && !generators) || !observable || !params || !out ||
        checkpoint_interval < 0)
        return TE_ERR_ARG;
    *out = nullptr;
    if (n_params > kAdjMaxLabels)
        return fail(ctx, TE_ERR_UNSUPPORTED, "at most %d generators", kAdjMaxLabels);
    if (H->n_rows != H->n_cols || H->n_rows == 0) return fail(ctx, TE_ERR_FATAL, "Invalid Hilbertspace dimension");
    if (H->n_rows > (size_t)INT32_MAX - 1024)                 // row indices of the node passes are 32-bit
        return fail(ctx, TE_ERR_UNSUPPORTED, "dimension above 2^31 - 1024");
    if (!observable->diagonal || observable->n_rows != H->n_rows)
        return fail(ctx, TE_ERR_UNSUPPORTED, "the loss observable must be a diagonal matrix of H's dimension");
    if (!(params->t > 0)) return fail(ctx, TE_ERR_ARG, "t must be > 0");
    if (params->m < 1) return fail(ctx, TE_ERR_ARG, "Krylov dimension must be >= 1");
    for (int p = 0; p < n_params; p++)
        if (!generators[p] || generators[p]->n_rows != H->n_rows || generators[p]->n_cols != H->n_cols)
            return fail(ctx, TE_ERR_ARG, "generator %d: null or wrong dimension", p);
    hipSetDevice(ctx->device);
    te_adjoint* a = new te_adjoint();
    a->ctx = ctx;
    a->H = H;
    a->O = observable;
    a->P = n_params;
    a->prm = *params;
    a->prm.sampling_step = params->t;          // samples at 0 and t only
    a->K = checkpoint_interval;
    const int64_t n = H->n_rows;
    a->n = n;
    a->m0 = (int)std::min<int64_t>(params->m, n);
    if (a->m0 > kMaxKrylov) {
        adj_free(a);
        return fail(ctx, TE_ERR_UNSUPPORTED, "Krylov dimension %d exceeds %d", a->m0, kMaxKrylov);
    }
    double nrm1 = 0;
    TE_TRY(te_matrix_norm1(H, &nrm1));
    a->tolRate = params->tol / params->t;
    a->numErrEst = (double)n * nrm1 * DBL_EPSILON;
    a->grid = stream_grid(ctx, n);
    a->pgrid = project_grid(ctx, n);
    if (H->use_vd && H->vd.perm) {
        a->perm = H->vd.perm;
        a->iperm = H->vd.iperm;
    }
    hipStream_t s = ctx->stream;

    // generators: off-diagonal entries labelled by generator, diagonal parts as dense rows,
    // all in the evolver's order (row r of the copy is row perm[r] of the basis order)
    std::vector<int> hperm(n), hiperm(n);
    if (a->perm) {
        TE_TRYH(hipMemcpy(hperm.data(), a->perm, sizeof(int) * n, hipMemcpyDeviceToHost));
        TE_TRYH(hipMemcpy(hiperm.data(), a->iperm, sizeof(int) * n, hipMemcpyDeviceToHost));
    } else {
        for (int64_t i = 0; i < n; i++) hperm[i] = hiperm[i] = (int)i;
    }
    std::vector<std::vector<int64_t>> gp(n_params);
    std::vector<std::vector<int>> gc(n_params);
    std::vector<std::vector<double2>> gv(n_params);
    std::vector<double> hdiag((size_t)n_params * n, 0.0);
    int64_t off_nnz = 0;
    for (int p = 0; p < n_params; p++) {
        te_matrix* G = generators[p];
        gp[p].resize(n + 1);
        gc[p].resize(G->nnz);
        gv[p].resize(G->nnz);
        TE_TRY(te_matrix_to_csr(G, (double*)gv[p].data(), gp[p].data(), gc[p].data()));
        for (int64_t r = 0; r < n; r++)
            for (int64_t k = gp[p][r]; k < gp[p][r + 1]; k++) {
                if (gc[p][k] == r) {
                    hdiag[(size_t)p * n + hiperm[r]] += gv[p][k].x;
                } else {
                    off_nnz++;
                }
            }
    }
    if (off_nnz >= ((int64_t)1 << 31)) {
        adj_free(a);
        return fail(ctx, TE_ERR_UNSUPPORTED, "too many generator entries");
    }
    a->g_nnz = off_nnz;
    a->off_label.assign(n_params, -1);
    for (int p = 0; p < n_params; p++) {
        bool has = false;
        for (int64_t r = 0; r < n && !has; r++)
            for (int64_t k = gp[p][r]; k < gp[p][r + 1]; k++)
                if (gc[p][k] != r) {
                    has = true;
                    break;
                }
        if (has) a->off_label[p] = a->n_off++;
    }
    if (off_nnz > 0) {
        // rows of the evolver order in slices of 64, slot-major, padded to the slice's longest row
        std::vector<int> len(n, 0);
        for (int64_t r = 0; r < n; r++) {
            const int ro = hperm[r];
            for (int p = 0; p < n_params; p++)
                for (int64_t e = gp[p][ro]; e < gp[p][ro + 1]; e++)
                    if (gc[p][e] != ro) len[r]++;
        }
        const int ns = (int)((n + 63) / 64);
        std::vector<int> hoff(ns + 1, 0);
        for (int sl = 0; sl < ns; sl++) {
            int mx = 0;
            for (int l = 0; l < 64 && (int64_t)sl * 64 + l < n; l++) mx = std::max(mx, len[(size_t)sl * 64 + l]);
            if (mx > 65535) {
                adj_free(a);
                return fail(ctx, TE_ERR_UNSUPPORTED, "generator rows longer than 65535 entries");
            }
            if ((int64_t)hoff[sl] + (int64_t)mx * 64 >= ((int64_t)1 << 31)) {
                adj_free(a);
                return fail(ctx, TE_ERR_UNSUPPORTED, "too many generator entries");
            }
            hoff[sl + 1] = hoff[sl] + mx * 64;
        }
        const size_t slots = (size_t)hoff[ns];
        std::vector<int> hcol(slots, 0);
        std::vector<unsigned char> hlab(slots, 0);
        std::vector<double2> hval(slots, make_double2(0.0, 0.0));
        std::vector<unsigned short> hlen(n);
        std::map<std::pair<unsigned long long, unsigned long long>, int> table;
        std::vector<unsigned short> hvid(slots, 0);
        std::vector<double2> htab;
        for (int64_t r = 0; r < n; r++) {
            const int ro = hperm[r];
            const int sl = (int)(r / 64), l = (int)(r % 64);
            int j = 0;
            for (int p = 0; p < n_params; p++)
                for (int64_t e = gp[p][ro]; e < gp[p][ro + 1]; e++)
                    if (gc[p][e] != ro) {
                        const size_t k = (size_t)hoff[sl] + (size_t)j * 64 + l;
                        hcol[k] = hiperm[gc[p][e]];
                        hlab[k] = (unsigned char)a->off_label[p];
                        hval[k] = gv[p][e];
                        unsigned long long kx, ky;
                        std::memcpy(&kx, &gv[p][e].x, 8);
                        std::memcpy(&ky, &gv[p][e].y, 8);
                        auto it = table.find({kx, ky});
                        if (it == table.end()) {
                            it = table.emplace(std::make_pair(kx, ky), (int)htab.size()).first;
                            htab.push_back(gv[p][e]);
                        }
                        hvid[k] = (unsigned short)std::min(it->second, 65535);
                        j++;
                    }
            hlen[r] = (unsigned short)j;
        }
        a->n_slices = ns;
        TE_TRY(aalloc(ctx, &a->soff, ns + 1));
        TE_TRY(aalloc(ctx, &a->scol, slots));
        TE_TRY(aalloc(ctx, &a->slab, slots));
        TE_TRY(aalloc(ctx, &a->slen, n));
        TE_TRYH(hipMemcpy(a->soff, hoff.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice));
        TE_TRYH(hipMemcpy(a->scol, hcol.data(), sizeof(int) * slots, hipMemcpyHostToDevice));
        TE_TRYH(hipMemcpy(a->slab, hlab.data(), slots, hipMemcpyHostToDevice));
        TE_TRYH(hipMemcpy(a->slen, hlen.data(), sizeof(unsigned short) * n, hipMemcpyHostToDevice));
        if ((int)htab.size() <= kTermMaxValues) {        // values through the table
            a->n_values = (int)htab.size();
            TE_TRY(aalloc(ctx, &a->svid, slots));
            TE_TRY(aalloc(ctx, &a->vtab, htab.size()));
            TE_TRYH(hipMemcpy(a->svid, hvid.data(), sizeof(unsigned short) * slots, hipMemcpyHostToDevice));
            TE_TRYH(hipMemcpy(a->vtab, htab.data(), sizeof(double2) * htab.size(), hipMemcpyHostToDevice));
        } else {
            TE_TRY(aalloc(ctx, &a->sval, slots));
            TE_TRYH(hipMemcpy(a->sval, hval.data(), sizeof(double2) * slots, hipMemcpyHostToDevice));
        }
    }
    TE_TRY(aalloc(ctx, &a->gdiag, (size_t)std::max(n_params, 1) * n));
    if (n_params > 0)
        TE_TRYH(hipMemcpy(a->gdiag, hdiag.data(), sizeof(double) * n_params * n, hipMemcpyHostToDevice));
    TE_TRY(aalloc(ctx, &a->odiag, n));
    if (a->perm) {
        launch_gather_re(observable->diag_re, a->perm, a->odiag, n, s);
    } else {
        TE_TRYH(hipMemcpyAsync(a->odiag, observable->diag_re, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    }

    // forward and backward state
    TE_TRY(aalloc(ctx, &a->psi0, n));
    TE_TRY(aalloc(ctx, &a->psiT, n));
    TE_TRY(aalloc(ctx, &a->scratch, n));
    if (a->K >= 1) TE_TRY(aalloc(ctx, &a->U, (size_t)n * (a->m0 + 1)));
    TE_TRY(aalloc(ctx, &a->ks, 1));
    TE_TRY(aalloc(ctx, &a->ks_hist, kAdjMaxSteps));
    TE_TRY(aalloc(ctx, &a->st, 1));
    TE_TRYH(hipHostMalloc((void**)&a->st_host, sizeof(StepState)));
    TE_TRY(aalloc(ctx, &a->wcoef, kMaxKrylov));
    TE_TRY(aalloc(ctx, &a->Y, 2 * kMaxKrylov));
    TE_TRY(aalloc(ctx, &a->trace, kAdjMaxSteps));
    TE_TRY(aalloc(ctx, &a->obs, 2));
    TE_TRY(upload_quadrature(ctx, a->quad));
    TE_TRY(reduce_scratch_alloc(ctx, a->rs,
                                std::max({a->grid, a->pgrid, H->shape.grid, H->vd.grid, H->vd.sell_grid})));
    TE_TRY(aalloc(ctx, &a->Vl, (size_t)n * (a->m0 + 1)));
    if (a->K >= 1) TE_TRY(aalloc(ctx, &a->Vp, (size_t)n * (a->m0 + 1)));
    TE_TRY(aalloc(ctx, &a->ksl, 1));
    TE_TRY(aalloc(ctx, &a->ksp, 1));
    TE_TRY(aalloc(ctx, &a->scale, 2));
    std::vector<double2> gln;
    std::vector<int> glo;
    gl_table(gln, glo);
    TE_TRY(aalloc(ctx, &a->gl, gln.size()));
    TE_TRY(aalloc(ctx, &a->gl_off, glo.size()));
    TE_TRYH(hipMemcpy(a->gl, gln.data(), sizeof(double2) * gln.size(), hipMemcpyHostToDevice));
    TE_TRYH(hipMemcpy(a->gl_off, glo.data(), sizeof(int) * glo.size(), hipMemcpyHostToDevice));
    TE_TRY(aalloc(ctx, &a->coef0, (size_t)kAdjMaxNodes * kMaxKrylov));
    TE_TRY(aalloc(ctx, &a->coef1, (size_t)kAdjMaxNodes * kMaxKrylov));
    TE_TRY(aalloc(ctx, &a->cend0, kMaxKrylov));
    TE_TRY(aalloc(ctx, &a->cend1, kMaxKrylov));
    TE_TRY(aalloc(ctx, &a->w, 2 * kAdjMaxNodes));   // two steps' weights (alternating)
    TE_TRYH(hipStreamCreateWithFlags(&a->s2, hipStreamNonBlocking));
    TE_TRYH(hipEventCreateWithFlags(&a->ev_lanczos, hipEventDisableTiming));
    TE_TRYH(hipEventCreateWithFlags(&a->ev_term, hipEventDisableTiming));
    TE_TRYH(hipHostMalloc((void**)&a->step_host, sizeof(AdjStep)));
    TE_TRY(aalloc(ctx, &a->nodes_psi, (size_t)kMatNodes * n));
    TE_TRY(aalloc(ctx, &a->nodes_lam, (size_t)kMatNodes * n));
    // in low-rank form a step has at most m node states: the odd steps' blocks need no more
    const int m_nodes = std::min(kMatNodes, (a->m0 + kNodeGroup - 1) / kNodeGroup * kNodeGroup);
    TE_TRY(aalloc(ctx, &a->nodes_psi1, (size_t)m_nodes * n));
    TE_TRY(aalloc(ctx, &a->nodes_lam1, (size_t)m_nodes * n));
    TE_TRY(aalloc(ctx, &a->psi_next, n));
    TE_TRY(aalloc(ctx, &a->lam_next, n));
    TE_TRY(aalloc(ctx, &a->zsum, n));
    TE_TRY(aalloc(ctx, &a->offacc, kAdjMaxLabels));
    TE_TRY(aalloc(ctx, &a->dgrad, kAdjMaxLabels));
    TE_TRY(aalloc(ctx, &a->partials, (size_t)a->grid * kAdjMaxLabels));
    TE_TRY(aalloc(ctx, &a->ticket, 1));
    TE_TRYH(hipMemsetAsync(a->ticket, 0, sizeof(unsigned int), s));
    TE_TRYH(hipStreamSynchronize(s));
    *out = a;
    return TE_OK;
}

extern "C" void te_adjoint_destroy(te_adjoint* a) { te::adj_free(a); }

#undef TE_TRY
#undef TE_TRYH

namespace te {

static LanczosArgs lanczos_args(te_adjoint* a, double2* U, KrylovScalars* ks, int m, double2 alpha) {
    LanczosArgs la{};
    la.n = a->n;
    la.row_block = a->H->row_block;
    la.n_row_blocks = a->H->n_row_blocks;
    la.ptr = a->H->row_ptr;
    la.col = a->H->col;
    la.val = a->H->val;
    la.U = U;
    la.ld = a->n;
    la.m = m;
    la.alpha = alpha;
    la.tolRate = a->tolRate;
    la.ks = ks;
    la.partials = a->rs.partials;
    la.ticket = a->rs.ticket;
    return la;
}

// One Lanczos recurrence of m iterations on U from U[:,0] (normalised by s_0 in ks).
static void lanczos(te_adjoint* a, double2* U, KrylovScalars* ks, int m, double2 alpha) {
    LanczosArgs la = lanczos_args(a, U, ks, m, alpha);
    for (int j = 0; j < m; j++) {
        la.j = j;
        launch_lanczos_apply(la, a->H, a->ctx->stream);
        launch_lanczos_project(la, a->pgrid, a->ctx->stream);
    }
}

// Two recurrences of m iterations with the same alpha in lockstep (U1 with ks1, U2 with ks2):
// each iteration's applies share one pass over H, their projections one launch. The second
// recurrence's block partials sit after the first's in the reduction scratch.
static void lanczos2(te_adjoint* a, double2* U1, KrylovScalars* ks1, double2* U2, KrylovScalars* ks2, int m,
                     double2 alpha) {
    LanczosArgs la = lanczos_args(a, U1, ks1, m, alpha);
    LanczosArgs lb = lanczos_args(a, U2, ks2, m, alpha);
    lb.partials = a->rs.partials + (size_t)a->rs.grid * 32;
    lb.ticket = a->rs.ticket + 1;
    // launch shapes of the paired kernels (blocks per CU; launch-shape sweeps:
    // TE_ADJ_APPLY2_BLOCKS_PER_CU, TE_ADJ_PROJECT2_BLOCKS_PER_CU). On the L = N = 12 ring, K = 1,
    // with the partials summed by separate finish launches: backward 41.8-41.9 ms a run at 2 and
    // 1, 41.3 at 2 and 2, 42.6-42.7 with 3 or 4 blocks per CU for the apply.
    static const int apply_per_cu = [] {
        const char* v = std::getenv("TE_ADJ_APPLY2_BLOCKS_PER_CU");
        return v ? std::max(0, std::atoi(v)) : 2;
    }();
    static const int project_per_cu = [] {
        const char* v = std::getenv("TE_ADJ_PROJECT2_BLOCKS_PER_CU");
        return v ? std::max(1, std::atoi(v)) : 2;
    }();
    const int cap = 32 * a->rs.grid;                  // each vector's partials region
    auto shape = [&](int per_cu, int64_t items) {
        const int64_t g = std::min<int64_t>((int64_t)a->ctx->num_cus * per_cu, std::max<int64_t>(items, 8));
        return (int)std::min<int64_t>(cap / 8 * 8, (g + 7) / 8 * 8);
    };
    const int agrid = apply_per_cu > 0 ? shape(apply_per_cu, ((int64_t)a->H->vd.n_slices + 3) / 4) : 0;
    const int pgrid = shape(project_per_cu, (a->n + kBlock - 1) / kBlock);
    for (int j = 0; j < m; j++) {
        la.j = lb.j = j;
        launch_lanczos_apply2(la, lb, a->H, agrid, a->ctx->stream);
        launch_lanczos_project2(la, lb, pgrid, a->ctx->stream);
    }
}

// Normalises the Lanczos start U[:,0]: ks->s[0] = 1/|U[:,0]|, *scale = |U[:,0]|.
static void start_norm(te_adjoint* a, const double2* v, KrylovScalars* ks, double* scale) {
    launch_nrm2sq(v, a->n, a->rs, 0, a->ctx->stream);
    hipLaunchKernelGGL(start_norm_kernel, dim3(1), dim3(1), 0, a->ctx->stream, a->rs.result, ks, scale);
}

static int forward(te_adjoint* a, te_result* R) {
    te_ctx* ctx = a->ctx;
    hipStream_t s = ctx->stream;
    const int64_t n = a->n;
    const double2 alpha = make_double2(0.0, -1.0 * a->prm.exp_factor);
    StepState init{};
    init.t_step = 1.0 / 10.0;
    init.m = a->m0;
    init.index_samples = 1;
    *a->st_host = init;
    KrylovScalars ks0{};
    ks0.s[0] = 1.0;
    std::vector<double2> e1(kMaxKrylov, make_double2(0.0, 0.0));
    e1[0] = make_double2(1.0, 0.0);
    TE_HIP(ctx, hipMemcpyAsync(a->ks, &ks0, sizeof(ks0), hipMemcpyHostToDevice, s));
    TE_HIP(ctx, hipMemcpyAsync(a->st, a->st_host, sizeof(StepState), hipMemcpyHostToDevice, s));
    TE_HIP(ctx, hipMemcpyAsync(a->Y, e1.data(), sizeof(double2) * kMaxKrylov, hipMemcpyHostToDevice, s));
    const long long n_samples = 2;
    StepConsts C = step_consts(a->prm, a->tolRate, a->numErrEst, n_samples, a->quad);
    C.Y = a->Y;
    C.wcoef = a->wcoef;
    C.trace = a->trace;
    C.trace_cap = kAdjMaxSteps;
    a->t_steps.clear();
    a->m_steps.clear();
    a->n_ckpt = 0;
    auto basis = [&](size_t k) -> double2* {
        if (a->K >= 1) return a->U;
        while (a->bases.size() <= k) {
            double2* p = nullptr;
            if (hipMalloc(&p, sizeof(double2) * (size_t)n * (a->m0 + 1)) != hipSuccess) return nullptr;
            a->bases.push_back(p);
        }
        return a->bases[k];
    };
    double2* U = basis(0);
    if (!U) return fail(ctx, TE_ERR_ALLOC, "no memory for the forward Krylov bases");
    TE_HIP(ctx, hipMemcpyAsync(U, a->psi0, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
    int m = a->m0;
    StepState& h = *a->st_host;
    int stalled = 0;
    CombineArgs cb{};
    cb.n = n;
    cb.ld = n;
    cb.wcoef = a->wcoef;
    cb.Y = a->Y;
    cb.diag = a->odiag;
    cb.nobs = 1;
    cb.obs_base = 0;
    cb.obs_values = a->obs;
    cb.n_samples = n_samples;
    cb.partials = a->rs.partials;
    // sample 0: <O> of psi0
    {
        CombineArgs c0 = cb;
        c0.U = U;
        c0.m = 1;
        c0.U0 = U;
        c0.write_w = 0;
        c0.y_base = 0;
        c0.S = 1;
        c0.write_last = 0;
        launch_combine(c0, a->grid, s);
    }
    size_t k = 0;
    while (h.index_samples < n_samples) {
        if (k >= (size_t)kAdjMaxSteps) return fail(ctx, TE_ERR_UNSUPPORTED, "more than %d Krylov steps", kAdjMaxSteps);
        if (a->K >= 1 && k % a->K == 0) {                 // checkpoint psi(t_k)
            if (a->n_ckpt == a->ckpt.size()) {
                double2* c = nullptr;
                TE_HIP(ctx, hipMalloc(&c, sizeof(double2) * n));
                a->ckpt.push_back(c);
            }
            TE_HIP(ctx, hipMemcpyAsync(a->ckpt[a->n_ckpt++], U, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
        }
        lanczos(a, U, a->ks, m, alpha);
        launch_step_control(C, a->st, a->ks, s);
        TE_HIP(ctx, hipGetLastError());
        TE_HIP(ctx, hipMemcpyAsync(a->st_host, a->st, sizeof(StepState), hipMemcpyDeviceToHost, s));
        TE_HIP(ctx, hipMemcpyAsync(a->ks_hist + k, a->ks, sizeof(KrylovScalars), hipMemcpyDeviceToDevice, s));
        TE_HIP(ctx, hipStreamSynchronize(s));
        if (h.fatal == 1)
            return fail(ctx, TE_ERR_FATAL, "Error: No small enough time step found to meet tolerance requirements.");
        if (h.fatal == 2) return fail(ctx, TE_ERR_FATAL, "Internal error: eigen-solver did not converge");
        stalled = (h.t_step == 0.0 && h.S == 0) ? stalled + 1 : 0;
        if (stalled >= 2) return fail(ctx, TE_ERR_FATAL, "time evolution makes no progress (t_step = 0)");
        m = h.m;
        a->t_steps.push_back(h.t_step);
        a->m_steps.push_back(h.m);
        const bool last = h.index_samples >= n_samples;
        double2* Unext = last ? a->scratch : basis(k + 1);
        if (!Unext) return fail(ctx, TE_ERR_ALLOC, "no memory for the forward Krylov bases");
        CombineArgs c1 = cb;
        c1.U = U;
        c1.m = m;
        c1.U0 = Unext;                 // currentVec -> column 0 of the next step's basis
        c1.write_w = h.write_w;
        c1.y_base = h.sample_base;
        c1.S = h.S;
        c1.write_last = h.S > 0 ? 1 : 0;
        c1.last_out = a->psiT;
        launch_combine(c1, a->grid, s);
        if (!h.write_w) {             // breakdown step: currentVec is the sampled state
            TE_HIP(ctx, hipMemcpyAsync(Unext, a->psiT, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
        }
        TE_HIP(ctx, hipGetLastError());
        U = Unext;
        k++;
    }
    // result and status (:445-504)
    launch_nrm2sq(a->psiT, n, a->rs, 0, s);
    double2 r2;
    TE_HIP(ctx, hipMemcpyAsync(&r2, a->rs.result, sizeof(double2), hipMemcpyDeviceToHost, s));
    TE_HIP(ctx, hipStreamSynchronize(s));
    R->err = h.err;
    R->evolved_time = h.t_now;
    R->num_samples = (size_t)h.index_samples;
    R->n_steps = (size_t)h.n_steps;
    R->dim = (size_t)n;
    R->krylov_dim = (size_t)h.m;
    R->status_code = final_status(h, a->numErrEst, a->prm.tol, std::sqrt(r2.x));
    return TE_OK;
}

static int backward(te_adjoint* a) {
    te_ctx* ctx = a->ctx;
    hipStream_t s = ctx->stream;
    // a previous run that ended early on an error may have left a term-wise pass queued there
    TE_HIP(ctx, hipStreamSynchronize(a->s2));
    const int64_t n = a->n;
    const double2 alpha_b = make_double2(0.0, 1.0 * a->prm.exp_factor);   // +i e H: evolution back in time
    const int nsteps = (int)a->t_steps.size();
    const int eg = std::max(8, std::min<int>((int)((n + 255) / 256), ctx->num_cus * 4)) / 8 * 8;
    hipLaunchKernelGGL(diag_mul_kernel, dim3(eg), dim3(256), 0, s, a->odiag, a->psiT, a->Vl, n);   // lambda(T)
    if (a->K >= 1)
        TE_HIP(ctx, hipMemcpyAsync(a->Vp, a->psiT, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
    TE_HIP(ctx, hipMemsetAsync(a->zsum, 0, sizeof(double) * n, s));
    TE_HIP(ctx, hipMemsetAsync(a->offacc, 0, sizeof(double) * kAdjMaxLabels, s));
    a->info.backward_err = 0;
    a->info.psi_err = 0;
    a->info.nodes = 0;
    a->info.node_states = 0;
    a->info.max_phase = 0;
    a->info.capped = 0;
    // launch shapes of the node passes (blocks per CU; launch-shape sweeps: TE_ADJ_MAT_BLOCKS_PER_CU,
    // TE_ADJ_TERM_BLOCKS_PER_CU, 0: the streaming kernels' grid, also the largest). On the
    // L = N = 12 ring, K = 1: 2 blocks per CU for the term-wise pass took 72.4 ms a run against
    // 74.0 ms with the streaming grid (4); 2, 4 or 8 for the materialisation within 0.1 ms.
    // node states on the matrix cores (materialise_mfma_kernel), or with TE_ADJ_MAT=valu (read at
    // every run) by the vector form
    const char* mat_env = std::getenv("TE_ADJ_MAT");
    const bool mat_mfma = !(mat_env && std::strcmp(mat_env, "valu") == 0);
    static const int mat_per_cu = [] {
        const char* v = std::getenv("TE_ADJ_MAT_BLOCKS_PER_CU");
        return v ? std::max(1, std::atoi(v)) : 8;
    }();
    static const int term_per_cu = [] {
        const char* v = std::getenv("TE_ADJ_TERM_BLOCKS_PER_CU");
        return v ? std::max(0, std::atoi(v)) : 2;
    }();
    const int mg = std::max(8, std::min<int>((int)((n + 255) / 256), ctx->num_cus * mat_per_cu)) / 8 * 8;
    const int tg = term_per_cu > 0            // at most a->grid: the size of the block partials
                       ? std::min(a->grid, std::max(8, std::min<int>((int)((n + 255) / 256),
                                                                     ctx->num_cus * term_per_cu)) / 8 * 8)
                       : a->grid;
    // Term-wise passes gather the node blocks of one group per entry (TE_ADJ_TERM_GROUPS=2: two;
    // on the L = N = 12 ring, K = 1: backward 34.0 ms with one, 34.7 ms with two).
    static const int term_groups = [] {
        const char* v = std::getenv("TE_ADJ_TERM_GROUPS");
        return (v && std::atoi(v) == 2) ? 2 : 1;
    }();
    // The node sums in low-rank form, exact to kAdjRankTol of their norm (adjoint_rank_kernel);
    // TE_ADJ_RANK_TOL sets another tolerance, 0 keeps the quadrature nodes themselves (read at
    // every run).
    // Eigen-decompositions of the coefficient kernel: bisection and twisted factorisations, QL
    // for eigenvalues closer than 1e-3 ||T|| (TE_ADJ_EIG=ql: QL always; read at every run).
    const char* eig_env = std::getenv("TE_ADJ_EIG");
    const int eig_fast = (eig_env && std::strcmp(eig_env, "ql") == 0) ? 0 : 1;
    double rank_tol = kAdjRankTol;
    if (const char* v = std::getenv("TE_ADJ_RANK_TOL")) rank_tol = std::max(0.0, std::atof(v));
    // One step record per step. In low-rank form a step has at most min(nodes, m) <= 64 node
    // states, one materialisation pass and one term-wise pass, whose kernels read the counts and
    // Krylov dimensions from the record: the backward is queued without a host round trip and the
    // records are read once at the end. The quadrature nodes themselves (up to kAdjMaxNodes) need
    // passes of 64 nodes, counted by the host after each step's coefficient kernel.
    if (a->steps_cap < nsteps) {
        if (a->steps) hipFree(a->steps);
        a->steps = nullptr;
        a->steps_cap = 0;
        TE_HIP(ctx, hipMalloc(&a->steps, sizeof(AdjStep) * nsteps));
        a->steps_cap = nsteps;
    }
    // In low-rank form, step k's term-wise pass (the last reader of its node states and weights)
    // runs on the second stream once step k - 1's Lanczos recurrences are done, beside step
    // k - 1's coefficient and rank kernels (one block each: the rest of the device would idle)
    // and materialisation. Node states and weights alternate between two buffers by step parity,
    // and step k - 2's coefficient kernels (the next writer of step k's buffers) wait for it.
    const bool overlap = rank_tol > 0.0;
    TermArgs pend{};
    bool pending = false;          // step k + 1's pass, to be queued on the second stream
    bool term_queued = false;      // a pass queued there that the main stream has not joined
    auto term_args = [&](AdjStep* step, const double* w, int node0, const double2* psi_nodes,
                         const double2* lam_nodes) {
        TermArgs ta{};
        ta.n = n;
        ta.n_slices = a->n_slices;
        ta.soff = a->soff;
        ta.scol = a->scol;
        ta.slab = a->slab;
        ta.svid = a->svid;
        ta.sval = a->sval;
        ta.slen = a->slen;
        ta.vtab = a->vtab;
        ta.n_values = a->n_values;
        ta.n_labels = a->n_off;
        ta.psi = psi_nodes;
        ta.lam = lam_nodes;
        ta.w = w;
        ta.step = step;
        ta.node0 = node0;
        ta.zsum = a->zsum;
        ta.partials = a->partials;
        ta.out = a->offacc;
        return ta;
    };
    for (int k = nsteps - 1; k >= 0; k--) {
        const double hstep = a->t_steps[k];
        AdjStep* step = a->steps + k;
        const bool odd = overlap && (k & 1);
        double* wk = a->w + (size_t)(odd ? kAdjMaxNodes : 0);
        double2* psi_nodes = odd ? a->nodes_psi1 : a->nodes_psi;
        double2* lam_nodes = odd ? a->nodes_lam1 : a->nodes_lam;
        // lambda: Lanczos of +i e H from lambda(t_{k+1}); with K >= 1 also psi's, in lockstep
        start_norm(a, a->Vl, a->ksl, a->scale + 1);
        if (a->K == 0) {
            lanczos(a, a->Vl, a->ksl, a->m0, alpha_b);
        } else {
            // psi(t_{k+1}): the forward checkpoint when one is kept there, else the state
            // back-evolved by the previous backward step
            if (k + 1 < nsteps && (k + 1) % a->K == 0)
                TE_HIP(ctx, hipMemcpyAsync(a->Vp, a->ckpt[(k + 1) / a->K], sizeof(double2) * n,
                                           hipMemcpyDeviceToDevice, s));
            start_norm(a, a->Vp, a->ksp, a->scale);
            lanczos2(a, a->Vl, a->ksl, a->Vp, a->ksp, a->m0, alpha_b);
        }
        if (term_queued) {            // step k + 2's pass: its buffers are written next
            TE_HIP(ctx, hipStreamWaitEvent(s, a->ev_term, 0));
            term_queued = false;
        }
        if (pending) {
            TE_HIP(ctx, hipEventRecord(a->ev_lanczos, s));
            TE_HIP(ctx, hipStreamWaitEvent(a->s2, a->ev_lanczos, 0));
            launch_termwise(pend, tg, term_groups, a->s2);
            TE_HIP(ctx, hipEventRecord(a->ev_term, a->s2));
            pending = false;
            term_queued = true;
        }
        AdjCoefArgs ca{};
        ca.ks[1] = a->ksl;
        ca.m[1] = a->m0;
        ca.m_from_ks[1] = 1;          // m_hbd when lambda's recurrence broke down
        ca.from_end[1] = 1;
        ca.scale[1] = a->scale + 1;
        const double2* Upsi;
        if (a->K == 0) {
            Upsi = a->bases[k];
            ca.ks[0] = a->ks_hist + k;
            ca.m[0] = a->m_steps[k];
            ca.from_end[0] = 0;
            ca.scale[0] = nullptr;
        } else {
            Upsi = a->Vp;
            ca.ks[0] = a->ksp;
            ca.m[0] = a->m0;
            ca.m_from_ks[0] = 1;
            ca.from_end[0] = 1;
            ca.scale[0] = a->scale;
        }
        ca.h = hstep;
        ca.gl = a->gl;
        ca.gl_off = a->gl_off;
        ca.coef[0] = a->coef0;
        ca.coef[1] = a->coef1;
        ca.coef_end[0] = a->cend0;
        ca.coef_end[1] = a->cend1;
        ca.w = wk;
        ca.out = step;
        ca.eig_fast = eig_fast;
        launch_adjoint_coef(ca, rank_tol, s);
        int n_nodes = kMatNodes;      // in low-rank form: one pass covers every node state
        if (rank_tol <= 0.0) {
            AdjStep st{};
            TE_HIP(ctx, hipMemcpyAsync(a->step_host, step, sizeof(AdjStep), hipMemcpyDeviceToHost, s));
            TE_HIP(ctx, hipStreamSynchronize(s));
            st = *a->step_host;
            if (st.capped == 2) return fail(ctx, TE_ERR_FATAL, "Internal error: eigen-solver did not converge");
            n_nodes = st.n_nodes;
        }
        // node states, 64 per pass, and the states at t_k
        for (int node0 = 0; node0 < n_nodes; node0 += kMatNodes) {
            MatArgs mp{Upsi, n, n, 0, a->coef0, a->cend0, step, node0, psi_nodes,
                       (node0 == 0 && a->K >= 1) ? a->psi_next : nullptr};
            MatArgs ml{a->Vl, n, n, 1, a->coef1, a->cend1, step, node0, lam_nodes,
                       node0 == 0 ? a->lam_next : nullptr};
            // m <= m0: a 40-row coefficient table (41.6 KB) lets the blocks share CUs with the
            // term-wise pass's
            if (mat_mfma && a->m0 <= 40) {
                hipLaunchKernelGGL(materialise_mfma_kernel<40>, dim3(mg), dim3(256), 0, s, mp);
                hipLaunchKernelGGL(materialise_mfma_kernel<40>, dim3(mg), dim3(256), 0, s, ml);
            } else if (mat_mfma) {
                hipLaunchKernelGGL(materialise_mfma_kernel<kMaxKrylov>, dim3(mg), dim3(256), 0, s, mp);
                hipLaunchKernelGGL(materialise_mfma_kernel<kMaxKrylov>, dim3(mg), dim3(256), 0, s, ml);
            } else if (a->m0 <= 40) {
                hipLaunchKernelGGL(materialise_nodes_kernel<40>, dim3(mg), dim3(256), 0, s, mp);
                hipLaunchKernelGGL(materialise_nodes_kernel<40>, dim3(mg), dim3(256), 0, s, ml);
            } else {
                hipLaunchKernelGGL(materialise_nodes_kernel<kMaxKrylov>, dim3(mg), dim3(256), 0, s, mp);
                hipLaunchKernelGGL(materialise_nodes_kernel<kMaxKrylov>, dim3(mg), dim3(256), 0, s, ml);
            }
            if (overlap) {
                pend = term_args(step, wk, node0, psi_nodes, lam_nodes);
                pending = true;
            } else {
                launch_termwise(term_args(step, wk, node0, psi_nodes, lam_nodes), tg, term_groups, s);
            }
        }
        TE_HIP(ctx, hipGetLastError());
        // the next (earlier) step starts from the states at t_k
        TE_HIP(ctx, hipMemcpyAsync(a->Vl, a->lam_next, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
        if (a->K >= 1)
            TE_HIP(ctx, hipMemcpyAsync(a->Vp, a->psi_next, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
    }
    if (term_queued) TE_HIP(ctx, hipStreamWaitEvent(s, a->ev_term, 0));
    if (pending) launch_termwise(pend, tg, term_groups, s);    // the last step's pass: nothing to overlap
    if (a->P > 0)
        hipLaunchKernelGGL(diag_grad_kernel, dim3(tg), dim3(256), 0, s, a->gdiag, a->zsum, n, a->P, a->partials,
                           a->ticket, a->dgrad);
    TE_HIP(ctx, hipGetLastError());
    std::vector<AdjStep> steps(nsteps);
    if (nsteps > 0) {
        TE_HIP(ctx, hipMemcpyAsync(steps.data(), a->steps, sizeof(AdjStep) * nsteps, hipMemcpyDeviceToHost, s));
        TE_HIP(ctx, hipStreamSynchronize(s));
    }
    for (const AdjStep& st : steps) {
        a->info.nodes += st.n_quad;
        a->info.node_states += st.n_nodes;
        a->info.max_phase = std::max(a->info.max_phase, st.phase);
        a->info.backward_err += st.err[1];
        if (a->K >= 1) a->info.psi_err += st.err[0];
        if (st.capped) a->info.capped++;
        if (st.capped == 2) return fail(ctx, TE_ERR_FATAL, "Internal error: eigen-solver did not converge");
    }
    return TE_OK;
}

}  // namespace te

extern "C" int te_adjoint_run(te_adjoint* a, const double* psi0, double* loss, double* grad, te_result* result,
                              te_adjoint_info* info) {
    using namespace te;
    if (!a || !psi0 || !loss || (a->P > 0 && !grad)) return TE_ERR_ARG;
    te_ctx* ctx = a->ctx;
    hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const int64_t n = a->n;
    TE_HIP(ctx, hipMemcpyAsync(a->scratch, psi0, sizeof(double2) * n, hipMemcpyHostToDevice, s));
    if (a->perm) launch_gather(a->scratch, a->perm, a->psi0, n, s);
    else TE_HIP(ctx, hipMemcpyAsync(a->psi0, a->scratch, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
    launch_nrm2sq(a->psi0, n, a->rs, 0, s);
    double2 r;
    TE_HIP(ctx, hipMemcpyAsync(&r, a->rs.result, sizeof(double2), hipMemcpyDeviceToHost, s));
    TE_HIP(ctx, hipStreamSynchronize(s));
    if (std::fabs(std::sqrt(r.x) - 1.0) > a->prm.tol) return fail(ctx, TE_ERR_FATAL, "Initial vector is not normalized");
    te_result R{};
    hipEvent_t e0, e1, e2;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventCreate(&e2);
    hipEventRecord(e0, s);
    int rc = forward(a, &R);
    hipEventRecord(e1, s);
    if (rc == TE_OK) rc = backward(a);
    hipEventRecord(e2, s);
    double hl[2] = {0, 0};
    std::vector<double> off(kAdjMaxLabels), dg(kAdjMaxLabels);
    if (rc == TE_OK) {
        TE_HIP(ctx, hipMemcpyAsync(hl, a->obs, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
        TE_HIP(ctx, hipMemcpyAsync(off.data(), a->offacc, sizeof(double) * kAdjMaxLabels, hipMemcpyDeviceToHost, s));
        TE_HIP(ctx, hipMemcpyAsync(dg.data(), a->dgrad, sizeof(double) * kAdjMaxLabels, hipMemcpyDeviceToHost, s));
    }
    TE_HIP(ctx, hipStreamSynchronize(s));
    float f_ms = 0, b_ms = 0;
    hipEventElapsedTime(&f_ms, e0, e1);
    hipEventElapsedTime(&b_ms, e1, e2);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    if (rc != TE_OK) return rc;
    *loss = hl[1];
    for (int p = 0; p < a->P; p++) {
        const double o = a->off_label[p] >= 0 ? off[a->off_label[p]] : 0.0;
        grad[p] = 2.0 * a->prm.exp_factor * (o + dg[p]);
    }
    a->info.forward_ms = f_ms;
    a->info.backward_ms = b_ms;
    a->info.n_steps = (int)a->t_steps.size();
    if (result) *result = R;
    if (info) *info = a->info;
    return TE_OK;
}

extern "C" int te_adjoint_set_hamiltonian(te_adjoint* a, te_matrix* H) {
    using namespace te;
    if (!a || !H) return TE_ERR_ARG;
    te_ctx* ctx = a->ctx;
    if (H->n_rows != a->n || H->n_cols != a->n) return fail(ctx, TE_ERR_ARG, "Hamiltonian of another dimension");
    const int* perm = (H->use_vd && H->vd.perm) ? H->vd.perm : nullptr;
    if ((perm == nullptr) != (a->perm == nullptr)) return fail(ctx, TE_ERR_ARG, "Hamiltonian held in another order");
    if (perm) {
        std::vector<int> p0(a->n), p1(a->n);
        TE_HIP(ctx, hipMemcpy(p0.data(), a->perm, sizeof(int) * a->n, hipMemcpyDeviceToHost));
        TE_HIP(ctx, hipMemcpy(p1.data(), perm, sizeof(int) * a->n, hipMemcpyDeviceToHost));
        if (p0 != p1) return fail(ctx, TE_ERR_ARG, "Hamiltonian held in another order");
    }
    double nrm1 = 0;
    const int rc = te_matrix_norm1(H, &nrm1);
    if (rc != TE_OK) return rc;
    if (std::max({H->shape.grid, H->vd.grid, H->vd.sell_grid}) > a->rs.grid)
        return fail(ctx, TE_ERR_ARG, "Hamiltonian needs a larger reduction grid");
    a->H = H;
    a->perm = perm;
    a->iperm = perm ? H->vd.iperm : nullptr;
    a->numErrEst = (double)a->n * nrm1 * DBL_EPSILON;
    return TE_OK;
}

extern "C" int te_adjoint_state(te_adjoint* a, double* psiT) {
    using namespace te;
    if (!a || !psiT) return TE_ERR_ARG;
    te_ctx* ctx = a->ctx;
    hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const double2* src = a->psiT;
    if (a->perm) {
        launch_gather(a->psiT, a->iperm, a->scratch, a->n, s);
        src = a->scratch;
    }
    TE_HIP(ctx, hipMemcpyAsync(psiT, src, sizeof(double2) * a->n, hipMemcpyDeviceToHost, s));
    TE_HIP(ctx, hipStreamSynchronize(s));
    return TE_OK;
}
