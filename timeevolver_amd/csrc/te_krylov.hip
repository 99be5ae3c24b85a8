// Krylov-space work of timeEvolve kept on the device (gfx950), compiled with
// -ffp-contract=off so that every scalar of the step control is rounded exactly as the
// reference's x86-64 build rounds it.
//
//  * step_control_kernel (one block): the m x m eigen-decomposition that replaces
//    LAPACKE_zhseqr (core/matrixDataTypes.h:135-156, called at
//    core/krylovTimeEvolver.cpp:365), the error-bound quadrature (integrateError, :584-618),
//    the step search (findMaximalStepSize, :153-254) and the sample schedule (:405-437),
//    producing the Krylov coefficient vectors of the new state and of every sample.
//  * combine_kernel (grid): currentVec = V w (:382), sampledState = V y (:421) and the
//    diagonal-observable expectations Re<psi|O|psi> = sum_r Re(O_rr) |psi_r|^2
//    (krylovSpMatrixObservable::expectation, core/krylovObservables.cpp:173-194), one pass over
//    the Krylov basis for a whole chunk of samples.
//
// Krylov exponential: the Lanczos matrix is anti-Hermitian tridiagonal, H[j][j] = a_j,
// H[j+1][j] = -H[j][j+1] = beta_j. With D = diag((-i)^j), D H D^{-1} = -i T, T real symmetric
// tridiagonal with T[j][j] = -Im a_j and off-diagonal beta_j, so
//   exp(x H) e1 = D^{-1} Q exp(-i x Lambda) Q^T e1,   T = Q Lambda Q^T,
// and the error integrand is |[exp(xH)e1]_{m-1}| = |sum_k p_k e^{-i x lambda_k}|,
// p_k = Q[m-1][k] Q[0][k]: O(m) per quadrature node instead of two m x m products.
#include "te_krylov.h"
#include "te_reduce.h"

#include <cfloat>

namespace te {

constexpr int kCtl = 1024;           // threads of the step controller (16 wavefronts)
constexpr int kPts = 1024;           // integrand points per evaluation batch
constexpr int kGrp = 8;              // lanes per integrand point
constexpr int kQB = 32;              // integrals per batch
constexpr int kTmpStride = 65;       // eigen-solver scratch tmp[i * 65 + k] (no LDS bank conflicts)

struct CtlShared {
    double Q[kMaxKrylov * kMaxKrylov];  // eigenvectors: column k at Q[k * 64 + row]
    double lam[kMaxKrylov];
    double d[kMaxKrylov], e[kMaxKrylov];
    double p[kMaxKrylov], g[kMaxKrylov], s[kMaxKrylov];
    double2 ph[kMaxKrylov];
    union {                            // integration points, or the eigen-solver's scratch
        struct {
            double xs[kPts], fx[kPts], wt[kPts];
        };
        double T[kMaxKrylov * kTmpStride];
    };
    // a batch of independent error integrals
    double qa[kQB], qb[kQB], qc[kQB], qd[kQB];
    double qI1[kQB], qL1[kQB], qerr[kQB];
    double qret[kQB], qfloor[kQB], qcap[kQB];
    int qmethod[kQB], qdone[kQB], qfin[kQB], qok[kQB];
    int act[kQB];
    int nact;
    double h;
    int m;
    int info;
    int eflag;
};

// sqrt(a^2 + b^2) without the scaling steps of hypot when no overflow or underflow can occur.
__device__ __forceinline__ double fast_hypot(double a, double b) {
    const double s = a * a + b * b;
    return (s > 1e-290 && s < 1e290) ? sqrt(s) : hypot(a, b);
}

// ---------------------------------------------------------------------------------------
// Implicit QL with Wilkinson shifts on a symmetric tridiagonal matrix (Golub & Van Loan,
// section 8.3), run by one wavefront. Lane k keeps d[k] and e[k] in registers; the rotation
// scalars are a uniform chain that reads d[i], d[i+1], e[i] with lane reads (no LDS round trip),
// the convergence test of all k at once is one ballot, and lane r applies each rotation to row r
// of Q (LDS). The rotation's norm and its reciprocal come from the reciprocal square root with
// one Newton step (within an ulp or two of the rounded values). d, e: LDS, n entries in,
// eigenvalues out in d.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double lane_read(double v, int k) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ int wave_tridiag_ql(int n, double* d, double* e, double* Q) {
    const int lane = threadIdx.x & 63;
    for (int k = 0; k < n; k++) Q[k * 64 + lane] = (lane == k) ? 1.0 : 0.0;
    if (n == 1) return 0;
    double dr = lane < n ? d[lane] : 0.0;
    double er = lane < n - 1 ? e[lane] : 0.0;
    int info = 0;
    for (int l = 0; l < n && !info; l++) {
        int iter = 0;
        while (true) {
            // first k >= l with |e_k| <= eps (|d_k| + |d_{k+1}|), or n - 1
            const double dn = __shfl_down(dr, 1, 64);
            const bool conv = (lane >= n - 1) || (fabs(er) <= DBL_EPSILON * (fabs(dr) + fabs(dn)));
            const unsigned long long mask = __ballot(conv) >> l;
            const int mm = l + __builtin_ctzll(mask);
            if (mm == l) break;
            if (++iter > 60) {
                info = 1;
                break;
            }
            const double dl = lane_read(dr, l), el = lane_read(er, l);
            double g = (lane_read(dr, l + 1) - dl) / (2.0 * el);
            double r = fast_hypot(g, 1.0);
            g = lane_read(dr, mm) - dl + el / (g + copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            bool early = false;
            for (int i = mm - 1; i >= l; i--) {
                const double ei = lane_read(er, i);
                const double f = s * ei;
                const double bb = c * ei;
                // r = |(f, g)| and 1/r: in the normal range from the hardware reciprocal square
                // root refined by one Newton step (no correctly rounded sqrt and divide in the
                // chain); hypot and a division otherwise
                const double ss = f * f + g * g;
                double ir;
                if (ss > 1e-290 && ss < 1e290) {
                    double y = __builtin_amdgcn_rsq(ss);
                    y = y * (1.5 - 0.5 * ss * y * y);
                    ir = y;
                    r = ss * y;
                } else {
                    r = hypot(f, g);
                    ir = r != 0.0 ? 1.0 / r : 0.0;
                }
                if (lane == i + 1) er = r;
                if (r == 0.0) {
                    const double di1 = lane_read(dr, i + 1) - p;
                    if (lane == i + 1) dr = di1;
                    if (lane == mm) er = 0.0;
                    early = true;
                    break;
                }
                s = f * ir;
                c = g * ir;
                g = lane_read(dr, i + 1) - p;
                r = (lane_read(dr, i) - g) * s + 2.0 * c * bb;
                p = s * r;
                if (lane == i + 1) dr = g + p;
                g = c * r - bb;
                if (lane < n) {
                    const double qf = Q[(i + 1) * 64 + lane];
                    const double qi = Q[i * 64 + lane];
                    Q[(i + 1) * 64 + lane] = s * qi + c * qf;
                    Q[i * 64 + lane] = c * qi - s * qf;
                }
            }
            if (early) continue;
            if (lane == l) {
                dr -= p;
                er = g;
            }
            if (lane == mm) er = 0.0;
        }
    }
    if (lane < n) d[lane] = dr;
    if (lane < n) e[lane] = er;
    return info;
}

// ---------------------------------------------------------------------------------------
// The same decomposition with one lane per eigenvalue, for matrices whose eigenvalues are
// apart: lane k finds the k-th smallest eigenvalue by bisection on the Sturm count (the number
// of negative pivots of the LDL^T factorisation of T - x I; Golub & Van Loan section 8.4, with
// the pivot guard of LAPACK's dlaebz), then its eigenvector from the twisted factorisation of
// T - lambda I: the top-down pivots D+ and bottom-up pivots D- meet at the index r where the twist
// element gamma_r = D+_r + D-_r - (d_r - lambda) is smallest, and the vector is built outward from
// x_r = 1 with the ratios -e_i / D+_i below r and -e_{i-1} / D-_i above (Dhillon and Parlett's
// construction). Every lane runs the same chains, so the wavefront's time is one eigenvalue's
// instead of the QL's sequence of rotations. An eigenvector built this way is accurate to about
// n eps ||T|| / gap; when two eigenvalues lie closer than kTwistGap ||T|| the routine returns 1
// without writing anything and the caller takes wave_tridiag_ql. On success: eigenvalues in d
// (ascending), eigenvector k in Q[k * 64 + row], e unchanged. tmp: 64 x 65 doubles of LDS.
// ---------------------------------------------------------------------------------------
constexpr double kTwistGap = 1e-3;

// 1 / q to within an ulp or two: the hardware reciprocal refined by two Newton steps.
__device__ __forceinline__ double recip(double q) {
    double r = __builtin_amdgcn_rcp(q);
    double t = fma(-q, r, 1.0);
    r = fma(r, t, r);
    t = fma(-q, r, 1.0);
    return fma(r, t, r);
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Gershgorin interval of T, ||T|| and the pivot guard of the Sturm recurrence (dlaebz: safe
// minimum x max(1, max e^2)), over the calling wavefront; bad: a non-finite entry.
struct TridiagBounds {
    double glo, ghi, tnorm, pivmin;
    bool bad;
};

__device__ TridiagBounds wave_tridiag_bounds(int n, const double* d, const double* e) {
    const int lane = threadIdx.x & 63;
    const bool own = lane < n;
    const bool bad = own && (!isfinite(d[lane]) || (lane + 1 < n && !isfinite(e[lane])));
    double glo = 1e300, ghi = -1e300, e2max = 0.0;
    if (own) {
        double rad = 0.0;
        if (lane > 0) rad += fabs(e[lane - 1]);
        if (lane + 1 < n) {
            rad += fabs(e[lane]);
            e2max = e[lane] * e[lane];
        }
        glo = d[lane] - rad;
        ghi = d[lane] + rad;
    }
    TridiagBounds b;
    b.bad = __ballot(bad) != 0;
    b.glo = wave_min(glo);
    b.ghi = wave_max(ghi);
    b.tnorm = fmax(fabs(b.glo), fabs(b.ghi));
    b.pivmin = DBL_MIN * fmax(1.0, wave_max(e2max));
    return b;
}

// Number of eigenvalues of T below x: negative pivots of the LDL^T factorisation of T - x I.
__device__ __forceinline__ int sturm_count(int n, const double* d, const double* e, double x, double pivmin) {
    double q = d[0] - x;
    if (fabs(q) <= pivmin) q = -pivmin;
    int c = q < 0.0;
    for (int i = 1; i < n; i++) {
        const double ei = e[i - 1];
        q = (d[i] - x) - ei * ei * recip(q);
        if (fabs(q) <= pivmin) q = -pivmin;
        c += q < 0.0;
    }
    return c;
}

// Lane k's eigenvector for its eigenvalue lam (lanes k < n, ascending), from the twisted
// factorisation; 1 (nothing written) when two eigenvalues lie closer than kTwistGap ||T||.
__device__ int wave_twisted_vectors(int n, double* d, const double* e, double* Q, double* tmp, double lam,
                                    double tnorm, double pivmin) {
    const int lane = threadIdx.x & 63;
    const bool own = lane < n;
    // gaps to the neighbouring eigenvalues
    const double up = __shfl_down(lam, 1, 64), dn = __shfl_up(lam, 1, 64);
    double gap = 1e300;
    if (own && lane + 1 < n) gap = fmin(gap, up - lam);
    if (own && lane > 0) gap = fmin(gap, lam - dn);
    if (wave_min(gap) < kTwistGap * tnorm) return 1;
    // bottom-up pivots D-_i into tmp, then top-down pivots D+_i into Q (as scratch, row i of
    // lane k at Q[i * 64 + k]) with the twist index r = argmin |gamma_i|
    int r = 0;
    if (own) {
        double dm = d[n - 1] - lam;
        if (fabs(dm) <= pivmin) dm = -pivmin;
        tmp[(n - 1) * kTmpStride + lane] = dm;
        for (int i = n - 2; i >= 0; i--) {
            const double ei = e[i];
            dm = (d[i] - lam) - ei * ei * recip(dm);
            if (fabs(dm) <= pivmin) dm = -pivmin;
            tmp[i * kTmpStride + lane] = dm;
        }
        double dp = d[0] - lam;
        if (fabs(dp) <= pivmin) dp = -pivmin;
        Q[lane] = dp;
        double gmin = fabs(tmp[lane]);                       // gamma_0 = D-_0
        for (int i = 1; i < n; i++) {
            const double ei = e[i - 1];
            dp = (d[i] - lam) - ei * ei * recip(dp);
            if (fabs(dp) <= pivmin) dp = -pivmin;
            Q[i * 64 + lane] = dp;
            const double g = fabs(dp + tmp[i * kTmpStride + lane] - (d[i] - lam));
            if (g < gmin) {
                gmin = g;
                r = i;
            }
        }
        // the vector, in tmp: x_r = 1, below r from D+ (Q scratch), above r from D- (tmp)
        double x = 1.0, s2 = 1.0;
        tmp[r * kTmpStride + lane] = 1.0;
        for (int i = r - 1; i >= 0; i--) {
            x = -e[i] * recip(Q[i * 64 + lane]) * x;
            tmp[i * kTmpStride + lane] = x;
            s2 += x * x;
        }
        x = 1.0;
        for (int i = r + 1; i < n; i++) {
            x = -e[i - 1] * recip(tmp[i * kTmpStride + lane]) * x;
            tmp[i * kTmpStride + lane] = x;
            s2 += x * x;
        }
        const double inv = 1.0 / sqrt(s2);
        for (int i = 0; i < n; i++) tmp[i * kTmpStride + lane] *= inv;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // Q[k * 64 + row] = component row of vector k (lane = row); rows past n stay zero
    for (int k = 0; k < n; k++) Q[k * 64 + lane] = own ? tmp[lane * kTmpStride + k] : 0.0;
    if (own) d[lane] = lam;
    return 0;
}

__device__ int wave_tridiag_bisect(int n, double* d, const double* e, double* Q, double* tmp) {
    const int lane = threadIdx.x & 63;
    const bool own = lane < n;
    // a non-finite entry goes to QL, which reports it through its iteration cap
    const TridiagBounds b = wave_tridiag_bounds(n, d, e);
    if (b.bad) return 1;
    if (n == 1) {
        if (lane == 0) Q[0] = 1.0;
        return 0;
    }
    // bisection: count(lo) <= k < count(hi) for k = lane
    double lo = b.glo - 2.0 * DBL_EPSILON * b.tnorm - b.pivmin, hi = b.ghi + 2.0 * DBL_EPSILON * b.tnorm + b.pivmin;
    // stop at 2 eps ||T|| (the accuracy the eigenvalues have anyway): at most about 55 halvings
    const double tol = 2.0 * DBL_EPSILON * b.tnorm + b.pivmin;
    bool active = own;
    int iter = 0;
    while (__ballot(active)) {
        // at most 2 + log2(range / tol) halvings; give up after 256 and leave the matrix to QL
        if (++iter > 256) return 1;
        const double mid = 0.5 * (lo + hi);
        if (active && (hi - lo <= tol || mid <= lo || mid >= hi)) active = false;
        if (__ballot(active) == 0) break;
        const int c = sturm_count(n, d, e, mid, b.pivmin);
        if (active) {
            if (c > lane) hi = mid;
            else lo = mid;
        }
    }
    return wave_twisted_vectors(n, d, e, Q, tmp, 0.5 * (lo + hi), b.tnorm, b.pivmin);
}

// The same decomposition by a whole block of W = blockDim.x / 64 wavefronts: each round every
// wavefront evaluates the Sturm count of every eigenvalue's interval at one of W inner points
// (lane k, point lo_k + (w + 1)(hi_k - lo_k)/(W + 1)), and the interval shrinks to the two
// neighbouring points that bracket the k-th eigenvalue: log2(W + 1) bits a round (16
// wavefronts: 4 bits) instead of one. Wavefront 0 then builds the vectors. Called by every
// thread of the block; returns the same value in all. scratch: LDS, 3 x 64 + W x 64 doubles
// (it may be the tmp of the vectors, which is written only after the last round).
__device__ int block_tridiag_multisect(int n, double* d, const double* e, double* Q, double* tmp, int* flag) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int W = blockDim.x >> 6;
    double* lo = tmp;                // 64
    double* hi = tmp + 64;           // 64
    double* cnt = tmp + 128;         // W x 64 counts (exact small integers)
    TridiagBounds b{};
    if (wv == 0) {
        b = wave_tridiag_bounds(n, d, e);
        if (lane < n) {
            lo[lane] = b.glo - 2.0 * DBL_EPSILON * b.tnorm - b.pivmin;
            hi[lane] = b.ghi + 2.0 * DBL_EPSILON * b.tnorm + b.pivmin;
        }
        if (lane == 0) flag[0] = b.bad ? 2 : (n == 1 ? 3 : 0);
    }
    __syncthreads();
    const int f0 = flag[0];
    if (f0 == 2) return 1;
    if (f0 == 3) {
        if (tid == 0) Q[0] = 1.0;
        return 0;
    }
    // every wavefront needs the guard and the tolerance: wavefront 0's values, through LDS
    if (wv == 0 && lane == 0) {
        cnt[W * 64] = b.tnorm;
        cnt[W * 64 + 1] = b.pivmin;
    }
    __syncthreads();
    const double tnorm = cnt[W * 64], pivmin = cnt[W * 64 + 1];
    const double tol = 2.0 * DBL_EPSILON * tnorm + pivmin;
    for (int round = 0;; round++) {
        if (round >= 64) return 1;                       // uniform: flag-driven exit below
        if (lane < n) {
            const double l = lo[lane], h = hi[lane];
            const double x = l + (double)(wv + 1) * ((h - l) / (double)(W + 1));
            cnt[wv * 64 + lane] = (h - l <= tol) ? 0.0 : (double)sturm_count(n, d, e, x, pivmin);
        }
        __syncthreads();
        if (wv == 0) {
            bool open = false;
            if (lane < n) {
                const double l = lo[lane], h = hi[lane];
                if (h - l > tol) {
                    double nl = l, nh = h;
                    for (int w = 0; w < W; w++) {
                        const double x = l + (double)(w + 1) * ((h - l) / (double)(W + 1));
                        if ((int)cnt[w * 64 + lane] <= lane) {
                            nl = x;
                        } else {
                            nh = x;
                            break;
                        }
                    }
                    lo[lane] = nl;
                    hi[lane] = nh;
                    open = nh - nl > tol && nl < nh;
                }
            }
            const bool any = __ballot(open) != 0;
            if (lane == 0) flag[0] = any ? 1 : 0;
        }
        __syncthreads();
        if (flag[0] == 0) break;
    }
    int info = 0;
    if (wv == 0) {
        const double lam = lane < n ? 0.5 * (lo[lane] + hi[lane]) : 0.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        info = wave_twisted_vectors(n, d, e, Q, tmp, lam, tnorm, pivmin);
        if (lane == 0) flag[0] = info;
    }
    __syncthreads();
    return flag[0];
}

// f(x) = h |sum_k p_k e^{-i x lambda_k}| at S.xs[0..npts) -> S.fx; kGrp lanes per point.
__device__ void eval_points(CtlShared& S, int npts) {
    const int tid = threadIdx.x;
    const int grp = tid / kGrp, q = tid % kGrp;
    const int m = S.m;
    for (int base = 0; base < npts; base += kCtl / kGrp) {
        const int pt = base + grp;
        double re = 0.0, im = 0.0;
        if (pt < npts) {
            const double x = S.xs[pt];
            for (int k = q; k < m; k += kGrp) {
                double sn, cs;
                sincos(x * S.lam[k], &sn, &cs);
                re += S.p[k] * cs;
                im -= S.p[k] * sn;
            }
        }
#pragma unroll
        for (int off = kGrp / 2; off > 0; off >>= 1) {
            re += __shfl_xor(re, off, kGrp);
            im += __shfl_xor(im, off, kGrp);
        }
        if (q == 0 && pt < npts) S.fx[pt] = S.h * hypot(re, im);
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------------
// Batched error integrals. The integrals q < nb over [qa[q], qb[q]] (method qmethod[q]: 0
// Gauss-Legendre 15, 1 Gauss-Legendre 7, 2 tanh-sinh) are independent, so their integrand
// points are evaluated together; each integral keeps exactly the summation order and the
// level schedule it would have alone. Results: qret[q] and qok[q] (integrateError's value and
// success flag, core/krylovTimeEvolver.cpp:584-618).
// ---------------------------------------------------------------------------------------

// Tanh-sinh (Takahasi & Mori) with Boost's driver structure: h = 1 on row 0, halving per row,
// error |I_k - I_{k-1}|, stop when error <= tol * L1, at most ts_rows - 1 refinements. Points
// that round onto an end point get weight 0.
__device__ void ts_batch(CtlShared& S, const StepConsts& C, double tol) {
    const double half_pi = 1.5707963267948966;
    const int tid = threadIdx.x;
    // level 0: centre, then (right, left) for nodes t = 1, 2, ...
    {
        const int n0 = C.ts_off[1] - C.ts_off[0];
        const int per = 1 + 2 * (n0 - 1);
        const int G = kPts / per;
        for (int g0 = 0; g0 < S.nact; g0 += G) {
            const int cnt = min(G, S.nact - g0);
            for (int i = tid; i < cnt * per; i += kCtl) {
                const int q = S.act[g0 + i / per], j = i % per;
                if (j == 0) {
                    S.xs[i] = S.qc[q];
                    S.wt[i] = 0.0;
                } else {
                    const double2 nd = C.ts_nodes[C.ts_off[0] + (j + 1) / 2];
                    const double off = S.qd[q] * nd.x;
                    const bool right = ((j - 1) & 1) == 0;
                    const double x = right ? S.qb[q] - off : S.qa[q] + off;
                    S.xs[i] = x;
                    S.wt[i] = (x != (right ? S.qb[q] : S.qa[q])) ? nd.y : 0.0;
                }
            }
            __syncthreads();
            eval_points(S, cnt * per);
            if (tid < cnt) {
                const int q = S.act[g0 + tid], base = tid * per;
                const double y0 = S.fx[base];
                double s = 0.0, as = 0.0;
                for (int j = 1; j < per; j++) {
                    s += S.fx[base + j] * S.wt[base + j];
                    as += fabs(S.fx[base + j]) * S.wt[base + j];
                }
                S.qI1[q] = (half_pi * y0 + s) * 1.0;
                S.qL1[q] = (half_pi * fabs(y0) + as) * 1.0;
                S.qerr[q] = DBL_MAX;
                S.qdone[q] = 0;
                S.qfin[q] = 1;
            }
            __syncthreads();
        }
    }
    double h = 1.0;
    for (int k = 1; k < C.ts_rows; k++) {
        if (tid == 0) {                                    // compact the still-active integrals
            int na = 0;
            for (int i = 0; i < S.nact; i++)
                if (!S.qdone[S.act[i]]) S.act[na++] = S.act[i];
            S.nact = na;
        }
        __syncthreads();
        const int nact = S.nact;
        if (nact == 0) break;
        h *= 0.5;
        const int off0 = C.ts_off[k];
        const int nk = C.ts_off[k + 1] - off0;
        const int per = 2 * nk;
        if (per <= kPts) {
            const int G = kPts / per;
            for (int g0 = 0; g0 < nact; g0 += G) {
                const int cnt = min(G, nact - g0);
                for (int i = tid; i < cnt * per; i += kCtl) {
                    const int q = S.act[g0 + i / per], j = i % per;
                    const double2 nd = C.ts_nodes[off0 + j / 2];
                    const double off = S.qd[q] * nd.x;
                    const bool right = (j & 1) == 0;
                    const double x = right ? S.qb[q] - off : S.qa[q] + off;
                    S.xs[i] = x;
                    S.wt[i] = (x != (right ? S.qb[q] : S.qa[q])) ? nd.y : 0.0;
                }
                __syncthreads();
                eval_points(S, cnt * per);
                if (tid < cnt) {
                    const int q = S.act[g0 + tid], base = tid * per;
                    double s = 0.0, as = 0.0;
                    for (int j = 0; j < per; j++) {
                        s += S.fx[base + j] * S.wt[base + j];
                        as += fabs(S.fx[base + j]) * S.wt[base + j];
                    }
                    const double I0 = S.qI1[q], L0 = S.qL1[q];
                    double I1 = 0.5 * I0, L11 = 0.5 * L0;
                    I1 += s * h;
                    L11 += as * h;
                    const double err = fabs(I0 - I1);
                    S.qI1[q] = I1;
                    S.qL1[q] = L11;
                    S.qerr[q] = err;
                    if (!isfinite(I1)) {
                        S.qfin[q] = 0;
                        S.qdone[q] = 1;
                    } else if (err <= tol * L11 ||
                               (err * S.qd[q] <= S.qfloor[q] && fabs(I1) * S.qd[q] <= S.qcap[q])) {
                        S.qdone[q] = 1;
                    }
                }
                __syncthreads();
            }
        } else {
            // rows wider than the point buffer: one integral at a time, points in chunks
            for (int ia = 0; ia < nact; ia++) {
                const int q = S.act[ia];
                double s = 0.0, as = 0.0;
                for (int done = 0; done < per; done += kPts) {
                    const int cnt = min(kPts, per - done);
                    for (int i = tid; i < cnt; i += kCtl) {
                        const int j = done + i;
                        const double2 nd = C.ts_nodes[off0 + j / 2];
                        const double off = S.qd[q] * nd.x;
                        const bool right = (j & 1) == 0;
                        const double x = right ? S.qb[q] - off : S.qa[q] + off;
                        S.xs[i] = x;
                        S.wt[i] = (x != (right ? S.qb[q] : S.qa[q])) ? nd.y : 0.0;
                    }
                    __syncthreads();
                    eval_points(S, cnt);
                    if (tid == 0)
                        for (int i = 0; i < cnt; i++) {
                            s += S.fx[i] * S.wt[i];
                            as += fabs(S.fx[i]) * S.wt[i];
                        }
                    __syncthreads();
                }
                if (tid == 0) {
                    const double I0 = S.qI1[q], L0 = S.qL1[q];
                    double I1 = 0.5 * I0, L11 = 0.5 * L0;
                    I1 += s * h;
                    L11 += as * h;
                    const double err = fabs(I0 - I1);
                    S.qI1[q] = I1;
                    S.qL1[q] = L11;
                    S.qerr[q] = err;
                    if (!isfinite(I1)) {
                        S.qfin[q] = 0;
                        S.qdone[q] = 1;
                    } else if (err <= tol * L11 ||
                               (err * S.qd[q] <= S.qfloor[q] && fabs(I1) * S.qd[q] <= S.qcap[q])) {
                        S.qdone[q] = 1;
                    }
                }
                __syncthreads();
            }
        }
    }
}

// Gauss-Legendre for the active integrals: centre first (odd n), then
// (f(c + d x_i) + f(c - d x_i)) w_i (Boost's gauss<double, N>::integrate order).
__device__ void gl_batch(CtlShared& S, const double2* nodes, int nn, int n) {
    const int tid = threadIdx.x;
    const int first = (n & 1) ? 1 : 0;
    const int per = first + 2 * (nn - first);
    const int G = kPts / per;
    for (int g0 = 0; g0 < S.nact; g0 += G) {
        const int cnt = min(G, S.nact - g0);
        for (int i = tid; i < cnt * per; i += kCtl) {
            const int q = S.act[g0 + i / per], j = i % per;
            if (first && j == 0) {
                S.xs[i] = S.qc[q];
            } else {
                const int node = first + (j - first) / 2;
                const double dx = S.qd[q] * nodes[node].x;
                S.xs[i] = ((j - first) & 1) == 0 ? S.qc[q] + dx : S.qc[q] - dx;
            }
        }
        __syncthreads();
        eval_points(S, cnt * per);
        if (tid < cnt) {
            const int q = S.act[g0 + tid], base = tid * per;
            double s = first ? S.fx[base] * nodes[0].y : 0.0;
            for (int i = first; i < nn; i++) {
                const int k = base + first + 2 * (i - first);
                s += (S.fx[k] + S.fx[k + 1]) * nodes[i].y;
            }
            S.qret[q] = s * S.qd[q];
            S.qok[q] = 1;
        }
        __syncthreads();
    }
}

// integrateError for the nb integrals set up in S.qa / S.qb / S.qmethod.
__device__ void integrate_batch(CtlShared& S, const StepConsts& C, int nb) {
    const double termination = 1e-3;                                 // :54
    const int tid = threadIdx.x;
    if (tid < nb) {
        S.qc[tid] = 0.5 * (S.qa[tid] + S.qb[tid]);
        S.qd[tid] = 0.5 * (S.qb[tid] - S.qa[tid]);
        // The integrand h |sum_k p_k e^{-i x lambda_k}| is bounded by h (sum_k |p_k| <= 1):
        // tanh-sinh refinement stops once successive estimates agree to eps h (b-a), the rounding
        // unit of that bound, provided the value is below termination (b-a) tolRate, where the
        // integral counts as successful whatever its error (core/krylovTimeEvolver.cpp:607-608).
        S.qfloor[tid] = DBL_EPSILON * S.h * (S.qb[tid] - S.qa[tid]);
        S.qcap[tid] = termination * (S.qb[tid] - S.qa[tid]) * C.tolRate;
    }
    for (int method = 0; method < 3; method++) {
        if (tid == 0) {
            int na = 0;
            for (int q = 0; q < nb; q++)
                if (S.qmethod[q] == method) S.act[na++] = q;
            S.nact = na;
        }
        __syncthreads();
        if (S.nact == 0) continue;
        if (method == 0) gl_batch(S, C.gl15, C.gl15_n, 15);
        else if (method == 1) gl_batch(S, C.gl7, C.gl7_n, 7);
        else {
            ts_batch(S, C, termination);
            if (tid < nb && S.qmethod[tid] == 2) {
                const int q = tid;
                const double d = S.qd[q];
                double ret = S.qI1[q] * d, error = S.qerr[q] * d;
                if (!S.qfin[q]) {                                    // Boost throws (:602-605)
                    ret = C.tol * 10;
                    error = 0.0;
                }
                S.qret[q] = ret;
                S.qok[q] = (error / ret > termination && ret > termination * (S.qb[q] - S.qa[q]) * C.tolRate) ? 0 : 1;
            }
            __syncthreads();
        }
    }
}

// out[j] = s_j [exp(x H) e1]_j  (expKrylov, :685-702, scaled for the unnormalised basis).
__device__ void exp_e1_scaled(CtlShared& S, double x, double2* out) {
    const int tid = threadIdx.x, m = S.m;
    if (tid < m) {
        double sn, cs;
        sincos(x * S.lam[tid], &sn, &cs);
        S.ph[tid] = make_double2(S.g[tid] * cs, -S.g[tid] * sn);
    }
    __syncthreads();
    if (tid < m) {
        double re = 0.0, im = 0.0;
        for (int k = 0; k < m; k++) {
            const double q = S.Q[k * 64 + tid];
            re += q * S.ph[k].x;
            im += q * S.ph[k].y;
        }
        double2 v;
        switch (tid & 3) {                    // multiply by i^j exactly
            case 0: v = make_double2(re, im); break;
            case 1: v = make_double2(-im, re); break;
            case 2: v = make_double2(-re, -im); break;
            default: v = make_double2(im, -re); break;
        }
        const double sj = S.s[tid];
        out[tid] = make_double2(v.x * sj, v.y * sj);
    }
    __syncthreads();
}

// Queue of precomputed integrals consumed in the reference's sequential order: `next` is the
// first unconsumed entry, `count` the number computed.
struct IntQueue {
    int next, count;
};

__device__ __forceinline__ void queue_set(CtlShared& S, int q, double a, double b, int method) {
    if (threadIdx.x == 0) {
        S.qa[q] = a;
        S.qb[q] = b;
        S.qmethod[q] = method;
    }
}

// findMaximalStepSize (core/krylovTimeEvolver.cpp:153-254). The decisions are the reference's,
// taken in its order; only the integrals are computed ahead in batches (speculative ones whose
// result the loop never consumes do not touch the success flag). Returns the status code, or -1
// where the reference exits after GO_MAX halvings.
__device__ int find_step(CtlShared& S, const StepConsts& C, double t_stepSuggestion, double t_step_max,
                         int n_substeps, bool increaseStep, int mLong, int mShort, double* t_stepRet,
                         double* err_stepRet) {
    const unsigned GO_MAX = 20;
    const double tolRate = C.tolRate;
    unsigned nbReductions = 0;
    double t_step = t_stepSuggestion;
    int code = 0;
    bool okInt = true;
    bool skipSubsteps = false;
    if (t_step > t_step_max) {
        t_step = t_step_max;
        skipSubsteps = true;
    }
    // batch 0: [0, t_step], then either the doubling candidates or the first substeps (valid if
    // no halving follows)
    int nb = 0;
    queue_set(S, nb++, 0, t_step, mLong);
    int spec_sub = 0;                       // substeps 1..spec_sub speculated with t_step unchanged
    // Speculation windows are kept small: a candidate far beyond the point where the
    // sequential loop stops (a doubling up to t_step_max) can cost more levels than every
    // consumed integral together.
    constexpr int kDoubleAhead = 2, kHalveAhead = 3;
    if (increaseStep && !skipSubsteps) {
        double v = t_step;
        while (nb < 1 + kDoubleAhead) {
            v *= 2;
            if (v > t_step_max) v = t_step_max;
            queue_set(S, nb++, 0, v, mLong);
            if (v >= t_step_max) break;
        }
    } else if (!skipSubsteps) {
        const double s = t_step / n_substeps;
        for (int n = 1; n <= 4 && nb < kQB; n++) {
            if (t_step + (n - 1) * s >= t_step_max) break;
            queue_set(S, nb++, t_step + (n - 1) * s, t_step + n * s, mShort);
            spec_sub = n;
        }
    }
    __syncthreads();
    integrate_batch(S, C, nb);
    IntQueue Q{1, nb};
    double err_step = S.qret[0];
    okInt = okInt && S.qok[0];

    if (increaseStep && !skipSubsteps) {
        double err_new = err_step, t_new = t_step;
        while (err_new < tolRate * t_new && t_step < t_step_max) {
            t_step = t_new;
            err_step = err_new;
            if (t_step >= t_step_max) {
                skipSubsteps = true;
                break;
            }
            t_new *= 2;
            if (t_new > t_step_max) t_new = t_step_max;
            if (Q.next >= Q.count) {                     // refill with the next doublings
                int k = 0;
                double v = t_new;
                queue_set(S, k++, 0, v, mLong);
                while (k < kDoubleAhead && v < t_step_max) {
                    v *= 2;
                    if (v > t_step_max) v = t_step_max;
                    queue_set(S, k++, 0, v, mLong);
                }
                __syncthreads();
                integrate_batch(S, C, k);
                Q.next = 0;
                Q.count = k;
            }
            err_new = S.qret[Q.next];
            okInt = okInt && S.qok[Q.next];
            Q.next++;
        }
        spec_sub = 0;
    }
    bool halved = false;
    if (err_step > tolRate * t_step) {
        halved = true;
        Q.next = Q.count = 0;
        while (err_step > tolRate * t_step) {
            nbReductions++;
            t_step = t_step / 2.0;
            if (Q.next >= Q.count) {                     // refill with the next halvings
                int k = 0;
                double v = t_step;
                while (k < kHalveAhead && nbReductions + k <= GO_MAX) {
                    queue_set(S, k++, 0, v, mLong);
                    v = v / 2.0;
                }
                __syncthreads();
                integrate_batch(S, C, k);
                Q.next = 0;
                Q.count = k;
            }
            err_step = S.qret[Q.next];
            okInt = okInt && S.qok[Q.next];
            Q.next++;
            if (nbReductions == GO_MAX) return -1;
            skipSubsteps = false;
        }
    }
    const double s = t_step / n_substeps;
    int n_s = 0;
    double deltaError = 0;
    if (!skipSubsteps) {
        // substeps speculated in batch 0 hold for the unchanged t_step only
        int sub_next = 1, sub_count = halved ? 0 : spec_sub;   // substep n is at queue slot offset
        int sub_slot0 = halved ? 0 : 1 + 0;                    // batch 0: slot 0 is [0, t_step]
        int batch = 4;
        while (err_step + deltaError < tolRate * (t_step + n_s * s)) {
            err_step += deltaError;
            n_s++;
            if (t_step + (n_s - 1) * s >= t_step_max) break;
            if (n_s > sub_count) {                       // refill: substeps n_s, n_s + 1, ...
                batch = min(kQB, batch * 2);
                int k = 0;
                for (int n = n_s; k < batch; n++) {
                    if (n > n_s && t_step + (n - 1) * s >= t_step_max) break;
                    queue_set(S, k++, t_step + (n - 1) * s, t_step + n * s, mShort);
                }
                __syncthreads();
                integrate_batch(S, C, k);
                sub_next = n_s;
                sub_count = n_s + k - 1;
                sub_slot0 = 0;
            }
            const int slot = sub_slot0 + (n_s - sub_next);
            deltaError = S.qret[slot];
            okInt = okInt && S.qok[slot];
        }
        t_step += (n_s - 1) * s;
    }
    if (err_step < C.numErrEst) {
        if (!skipSubsteps || err_step > 0.1 * tolRate * t_step) code += 1;
    }
    if (!okInt) code += 20;
    *t_stepRet = t_step;
    *err_stepRet = err_step;
    exp_e1_scaled(S, t_step, C.wcoef);
    return code;
}

// One Krylov step of timeEvolve after the Lanczos recurrence (:342-441), run by a block of
// kCtl threads. Returns 0, or the fatal code it also stores in st->fatal (1: no small enough
// step, :203-207; 2: the eigen-solver failed, :366-370); the value is the same in every thread.
__device__ int step_control_body(CtlShared& S, const StepConsts& C, StepState* st, KrylovScalars* ks) {
    const int tid = threadIdx.x;
    const long long clk0 = wall_clock64();
    StepState L = *st;
    const bool hbd = ks->broken != 0;
    int m = L.m;
    if (hbd) m = ks->m_hbd;
    L.n_steps += 1;
    L.breakdown = hbd ? 1 : 0;
    double t_step = L.t_step;
    if (hbd) {                                        // :342-363
        t_step = C.t - L.t_now;
        L.m = m;
        L.statusCode = 1;
    }
    if (tid < m) {
        S.d[tid] = -ks->a[tid].y;
        S.e[tid] = (tid + 1 < m) ? ks->beta[tid] : 0.0;
        S.s[tid] = ks->s[tid];
    }
    if (tid == 0) {
        S.m = m;
        S.h = ks->h;
    }
    __syncthreads();
    // multisection over the block's wavefronts and twisted factorisations, QL when eigenvalues
    // lie close (or on request)
    int einfo = C.eig_fast ? block_tridiag_multisect(m, S.d, S.e, S.Q, S.T, &S.eflag) : 1;
    if (tid < 64) {
        if (einfo) einfo = wave_tridiag_ql(m, S.d, S.e, S.Q);
        if (tid == 0) S.info = einfo;
    }
    __syncthreads();
    if (S.info != 0) {                               // :366-370
        if (tid == 0) st->fatal = 2;
        return 2;
    }
    if (tid < m) {
        S.lam[tid] = S.d[tid];
        S.g[tid] = S.Q[tid * 64];
        S.p[tid] = S.Q[tid * 64 + m - 1] * S.Q[tid * 64];
    }
    __syncthreads();
    const long long clk1 = wall_clock64();

    double err_step = 0;
    if (!hbd) {                                       // :374-383
        const int mLong = C.fastIntegration ? 0 : 2, mShort = C.fastIntegration ? 1 : 2;   // :269-277
        const double s0 = 0.97 * t_step;
        const int code = find_step(S, C, s0, C.t - L.t_now, 50, L.t_now == 0, mLong, mShort, &t_step, &err_step);
        if (code < 0) {
            if (tid == 0) st->fatal = 1;
            return 1;
        }
        L.errorCodeFindSubstep = code;
    }
    switch (L.errorCodeFindSubstep) {                 // :386-399
        case 0: break;
        case 1: L.nbErrRoundOff++; break;
        case 20: L.nbErrInt++; break;
        case 21: L.nbErrInt++; L.nbErrRoundOff++; break;
    }
    if (tid == 0 && L.n_steps - 1 < C.trace_cap) C.trace[L.n_steps - 1] = t_step;

    const long long clk2 = wall_clock64();
    L.sample_base = L.index_samples;                  // :405-437
    L.S = 0;
    while (L.index_samples < C.n_samples &&
           (L.t_sampling + C.samplingStep <= L.t_now + t_step ||
            (L.index_samples == C.n_samples - 1 && fabs(L.t_now + t_step - C.t) <= DBL_EPSILON))) {
        L.t_sampling += C.samplingStep;
        exp_e1_scaled(S, L.t_sampling - L.t_now, C.Y + (size_t)L.index_samples * kMaxKrylov);
        L.index_samples++;
        L.S++;
    }
    L.t_now_prev = L.t_now;
    L.err_prev = L.err;
    L.t_now += t_step;                                // :440-441
    L.err += err_step;
    L.t_step = t_step;
    L.write_w = hbd ? 0 : 1;
    const long long clk3 = wall_clock64();
    L.phase_ticks[0] += clk1 - clk0;
    L.phase_ticks[1] += clk2 - clk1;
    L.phase_ticks[2] += clk3 - clk2;
    __syncthreads();
    if (tid == 0) {
        *st = L;
        ks->broken = 0;
        ks->s[0] = 1.0;
    }
    return 0;
}

__global__ __launch_bounds__(kCtl) void step_control_kernel(StepConsts C, StepState* st, KrylovScalars* ks) {
    __shared__ CtlShared S;
    step_control_body(S, C, st, ks);
}

// Batched form: block i runs the Krylov step of instance i when it is live, then records what
// the round's combine pass needs and whether the instance goes on. The stall rule is the one
// the single-instance host loop applies: two consecutive steps with t_step = 0 and no sample.
__global__ __launch_bounds__(kCtl) void batch_step_control_kernel(const StepConsts* Cs, StepState* st,
                                                                  KrylovScalars* ks, BatchCtl* bc,
                                                                  long long n_samples) {
    __shared__ CtlShared S;
    const int i = blockIdx.x;
    if (!bc[i].live) {
        if (threadIdx.x == 0) bc[i].stepped = 0;
        return;
    }
    const int rc = step_control_body(S, Cs[i], st + i, ks + i);
    if (threadIdx.x == 0) {
        BatchCtl b = bc[i];
        if (rc != 0) {
            b.live = 0;
            b.stepped = 0;
        } else {
            const StepState L = st[i];
            b.stepped = 1;
            b.cm = L.m;
            b.stall = (L.t_step == 0.0 && L.S == 0) ? b.stall + 1 : 0;
            if (b.stall >= 2) {
                st[i].fatal = 3;
                b.live = 0;
            } else {
                b.live = L.index_samples < n_samples ? 1 : 0;
            }
        }
        bc[i] = b;
    }
}

// ---------------------------------------------------------------------------------------
// Adjoint: both Krylov spaces of a backward step decomposed at once (wavefront 0: basis 0,
// wavefront 1: basis 1), then the quadrature rule and every node's coefficient vectors.
//
// With sigma = -1 for a basis that starts at t_k and +1 for one that starts at t_{k+1}, basis
// b's vector at s = t_k + x varies as exp(i x sigma_b theta) over its Ritz values theta, so the
// integrand <lambda|G|psi> holds the frequencies sigma_0 theta_0 - sigma_1 theta_1. The
// Gauss-Legendre rule is chosen for the largest of them: the error bound of an n-point rule on
// exp(i omega x) over a length-H interval, relative to H, is
//   (omega H)^{2n} (n!)^4 / ((2n+1) ((2n)!)^3),
// and the rule is the fewest nodes (order n <= kGlMax times equal sub-intervals) bringing it
// below 1e-13.
// ---------------------------------------------------------------------------------------
struct AdjShared {
    double Q[2][kMaxKrylov * kMaxKrylov];
    double T[2][kMaxKrylov * kTmpStride];      // scratch of wave_tridiag_bisect
    double lam[2][kMaxKrylov], d[2][kMaxKrylov], e[2][kMaxKrylov], g[2][kMaxKrylov], s[2][kMaxKrylov];
    double p[2][kMaxKrylov];
    double lo[2], hi[2];
    int info[2];
    int mm[2];
    int order, n_sub;
    double phase;
};

// out[j] = scale s_j [exp(x H_m) e1]_j from the eigen-decomposition of basis b, computed by the
// 64 lanes of the calling wavefront (lane k forms the phase term of Ritz value k, shuffles carry
// it to the lanes that sum row j).
__device__ void adj_coef_wave(const AdjShared& S, int b, int m, double x, double scale, double2* out) {
    const int lane = threadIdx.x & 63;
    double phr = 0.0, phi = 0.0;
    if (lane < m) {
        double sn, cs;
        sincos(x * S.lam[b][lane], &sn, &cs);
        phr = S.g[b][lane] * cs;
        phi = -S.g[b][lane] * sn;
    }
    double re = 0.0, im = 0.0;
    for (int k = 0; k < m; k++) {
        const double pr = __shfl(phr, k, 64), pi = __shfl(phi, k, 64);
        if (lane < m) {
            const double q = S.Q[b][k * 64 + lane];
            re += q * pr;
            im += q * pi;
        }
    }
    if (lane < m) {
        double2 v;
        switch (lane & 3) {                    // multiply by i^j exactly
            case 0: v = make_double2(re, im); break;
            case 1: v = make_double2(-im, re); break;
            case 2: v = make_double2(-re, -im); break;
            default: v = make_double2(im, -re); break;
        }
        const double f = S.s[b][lane] * scale;
        out[lane] = make_double2(v.x * f, v.y * f);
    }
}

__device__ double gl_log_bound(int n, double phase) {
    if (phase <= 0.0) return -1e300;
    return 2.0 * n * log(phase) + 4.0 * lgamma(n + 1.0) - log(2.0 * n + 1.0) - 3.0 * lgamma(2.0 * n + 1.0);
}

__global__ __launch_bounds__(256) void adjoint_coef_kernel(AdjCoefArgs a) {
    __shared__ AdjShared S;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    if (tid < 2) {
        int m = a.m[tid];
        if (a.m_from_ks[tid] && a.ks[tid]->broken) m = a.ks[tid]->m_hbd;
        S.mm[tid] = m;
    }
    __syncthreads();
    if (wv < 2 && lane < S.mm[wv]) {
        const int b = wv, m = S.mm[b];
        const KrylovScalars* ks = a.ks[b];
        S.d[b][lane] = -ks->a[lane].y;
        S.e[b][lane] = (lane + 1 < m) ? ks->beta[lane] : 0.0;
        S.s[b][lane] = ks->s[lane];
    }
    __syncthreads();
    if (wv < 2) {                     // each wavefront decomposes one basis (no barrier inside)
        int info = a.eig_fast ? wave_tridiag_bisect(S.mm[wv], S.d[wv], S.e[wv], S.Q[wv], S.T[wv]) : 1;
        if (info) info = wave_tridiag_ql(S.mm[wv], S.d[wv], S.e[wv], S.Q[wv]);
        if (lane == 0) S.info[wv] = info;
    }
    __syncthreads();
    if (wv < 2 && lane < S.mm[wv]) {
        const int b = wv, m = S.mm[b];
        S.lam[b][lane] = S.d[b][lane];
        S.g[b][lane] = S.Q[b][lane * 64];
        S.p[b][lane] = S.Q[b][lane * 64 + m - 1] * S.Q[b][lane * 64];
    }
    __syncthreads();
    if (tid < 2) {
        const int b = tid;
        double lo = 1e300, hi = -1e300;
        for (int k = 0; k < S.mm[b]; k++) {
            const double v = (a.from_end[b] ? 1.0 : -1.0) * S.lam[b][k];
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
        S.lo[b] = lo;
        S.hi[b] = hi;
    }
    __syncthreads();
    if (tid == 0) {
        const double om = fmax(fabs(S.hi[0] - S.lo[1]), fabs(S.lo[0] - S.hi[1]));
        const double phase = om * a.h;
        const double lim = log(1e-13);
        int best_n = kGlMax, best_sub = kAdjMaxNodes / kGlMax, capped = 1;
        for (int sub = 1; sub * 2 <= kAdjMaxNodes; sub++) {
            int n = 2;
            while (n <= kGlMax && gl_log_bound(n, phase / sub) > lim) n++;
            if (n <= kGlMax && n * sub <= kAdjMaxNodes) {
                if (capped || n * sub < best_n * best_sub) {
                    best_n = n;
                    best_sub = sub;
                    capped = 0;
                }
                if (n <= kGlMax / 2) break;       // more sub-intervals only add nodes
            }
        }
        S.order = best_n;
        S.n_sub = best_sub;
        S.phase = phase;
        AdjStep o{};
        o.n_nodes = best_n * best_sub;
        o.n_quad = o.n_nodes;
        o.m[0] = S.mm[0];
        o.m[1] = S.mm[1];
        o.order = best_n;
        o.n_sub = best_sub;
        // 2: an eigen-decomposition failed; 1: the rule was capped at kAdjMaxNodes nodes
        o.capped = (S.info[0] != 0 || S.info[1] != 0) ? 2 : capped;
        o.phase = phase;
        *a.out = o;
    }
    __syncthreads();
    const int order = S.order, nsub = S.n_sub, nn = order * nsub;
    const double sl = a.h / nsub;
    // quadrature nodes and weights on [0, h]; node q = sub-interval q / order, point q % order
    for (int q = tid; q < nn; q += blockDim.x) {
        const double2 gw = a.gl[a.gl_off[order] + q % order];
        a.w[q] = 0.5 * sl * gw.y;
    }
    // coefficients: wavefront wv handles (node, basis) pairs wv, wv + 4, ...; the last pair of
    // each basis is its combination at t_k
    for (int it = wv; it < 2 * (nn + 1); it += 4) {
        const int b = it / (nn + 1), q = it % (nn + 1);
        double x;
        double2* out;
        if (q < nn) {
            const double2 gw = a.gl[a.gl_off[order] + q % order];
            x = sl * ((q / order) + 0.5 * (1.0 + gw.x));
            out = a.coef[b] + (size_t)q * kMaxKrylov;
        } else {
            x = 0.0;
            out = a.coef_end[b];
        }
        const double arg = a.from_end[b] ? a.h - x : x;
        adj_coef_wave(S, b, S.mm[b], arg, a.scale[b] ? *a.scale[b] : 1.0, out);
    }
    // error integrals h_b int |[exp(y H_m) e1]_{m-1}| over the step, with the same rule
    __syncthreads();
    if (wv < 2) {
        const int b = wv, m = S.mm[b];
        double acc = 0.0;
        for (int q = 0; q < nn; q++) {
            const double2 gw = a.gl[a.gl_off[order] + q % order];
            const double x = sl * ((q / order) + 0.5 * (1.0 + gw.x));
            const double arg = a.from_end[b] ? a.h - x : x;
            double re = 0.0, im = 0.0;
            if (lane < m) {
                double sn, cs;
                sincos(arg * S.lam[b][lane], &sn, &cs);
                re = S.p[b][lane] * cs;
                im = -S.p[b][lane] * sn;
            }
            re = wave_sum(re);
            im = wave_sum(im);
            acc += 0.5 * sl * gw.y * hypot(re, im);
        }
        if (lane == 0) a.out->err[b] = a.ks[b]->h * acc * (a.scale[b] ? *a.scale[b] : 1.0);
    }
}

// Low-rank form of a step's node sum. The gradient passes need, for the entries (r, c) of the
// generators, M[r][c] = sum_q w_q conj(lambda_q[r]) psi_q[c] with lambda_q = U_1 a_q and
// psi_q = U_0 b_q (a_q = coef[1] row q, b_q = coef[0] row q), that is M = conj(U_1) C U_0^T
// with the m_1 x m_0 matrix C[i][j] = sum_q w_q conj(a_q[i]) b_q[j]. Gram-Schmidt with column
// pivoting writes C = sum_k q_k R_k (q_k a unit column, R_k a row) and stops once the Frobenius
// norm of what is left is at most rank_tol ||C||_F, or after min(nodes, m_0, m_1) terms, when
// nothing is left (C has at most that rank). The r terms replace the nodes: a'_k = conj(q_k),
// b'_k = R_k, w'_k = 1, and n_nodes becomes r; n_quad keeps the rule's node count. Every later
// pass reads only n_nodes rows, so the node states and term-wise groups shrink from the
// rule's count to C's numerical rank: the node vectors vary smoothly over the step, so C's
// singular values fall geometrically (on the L = 10 ring, a step of length 1 with m = 40 and
// 28 nodes: 1e-14 of the largest at the 15th). The residual left out is exact, not
// estimated: each column is updated in place and its norm recomputed.
// One block; the Gram-Schmidt runs in wavefront 0, lane j owning column j of C (in LDS).
constexpr int kRankChunk = 8;      // nodes staged in LDS at a time while C is formed

// MS: the largest Krylov dimension the launch serves (the LDS stride of C and the staged
// nodes): 40 keeps the block at 36 KB of LDS, which fits beside the term-wise pass's blocks.
template <int MS>
__global__ __launch_bounds__(256) void adjoint_rank_kernel(AdjCoefArgs a, double rank_tol) {
    __shared__ double2 C[MS * MS];                       // C[i * MS + j]
    __shared__ double2 As[kRankChunk * MS], Bs[kRankChunk * MS];
    __shared__ double2 qv[MS];
    __shared__ int s_go, s_piv;
    __shared__ double s_inv;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int nn = a.out->n_quad;
    const int m0 = a.out->m[0], m1 = a.out->m[1];
    const double2* A = a.coef[1];
    const double2* B = a.coef[0];
    // C = sum_q w_q conj(a_q) b_q^T: thread tid accumulates the entries tid + 256 t (row i =
    // e / MS, column j = e % MS) over chunks of nodes staged in LDS (w_q folded into a_q)
    constexpr int kPer = (MS * MS + 255) / 256;
    double2 acc[kPer];
#pragma unroll
    for (int t = 0; t < kPer; t++) acc[t] = make_double2(0.0, 0.0);
    for (int q0 = 0; q0 < nn; q0 += kRankChunk) {
        const int nq = min(kRankChunk, nn - q0);
        for (int e = tid; e < nq * MS; e += blockDim.x) {
            const int q = e / MS, i = e % MS;
            const double w = a.w[q0 + q];
            const double2 x = i < m1 ? A[(size_t)(q0 + q) * kMaxKrylov + i] : make_double2(0.0, 0.0);
            As[e] = make_double2(w * x.x, w * x.y);
            Bs[e] = i < m0 ? B[(size_t)(q0 + q) * kMaxKrylov + i] : make_double2(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kPer; t++) {
            const int e = tid + 256 * t, i = e / MS, j = e % MS;
            if (e >= MS * MS) break;
            double re = acc[t].x, im = acc[t].y;
            for (int q = 0; q < nq; q++) {
                const double2 x = As[q * MS + i], y = Bs[q * MS + j];
                re += x.x * y.x + x.y * y.y;          // conj(x) y
                im += x.x * y.y - x.y * y.x;
            }
            acc[t] = make_double2(re, im);
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < kPer; t++)
        if (tid + 256 * t < MS * MS) C[tid + 256 * t] = acc[t];
    __syncthreads();
    const int rmax = min(nn, min(m0, m1));
    double thr = 0.0;
    int r = 0;
    for (;;) {
        if (wv == 0) {
            double nrm = 0.0;
            if (lane < m0)
                for (int i = 0; i < m1; i++) {
                    const double2 c = C[i * MS + lane];
                    nrm += c.x * c.x + c.y * c.y;
                }
            const double tot = wave_sum(nrm);
            if (r == 0) thr = rank_tol * rank_tol * tot;
            double best = lane < m0 ? nrm : -1.0;
            int bi = lane;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
            if (lane == 0) {
                s_go = r < rmax && (r == 0 || tot > thr);
                s_piv = bi;
                s_inv = best > 0.0 ? 1.0 / sqrt(best) : 0.0;
            }
        }
        __syncthreads();
        if (!s_go) break;
        if (wv == 0 && lane < m1) {
            const double2 c = C[lane * MS + s_piv];
            qv[lane] = make_double2(c.x * s_inv, c.y * s_inv);
        }
        __syncthreads();
        if (wv == 0) {
            if (lane < m0) {
                double rr = 0.0, ri = 0.0;
                for (int i = 0; i < m1; i++) {         // R = <q, C[:, lane]>
                    const double2 q = qv[i], c = C[i * MS + lane];
                    rr += q.x * c.x + q.y * c.y;
                    ri += q.x * c.y - q.y * c.x;
                }
                for (int i = 0; i < m1; i++) {         // C[:, lane] -= q R
                    const double2 q = qv[i];
                    double2& c = C[i * MS + lane];
                    c.x -= q.x * rr - q.y * ri;
                    c.y -= q.x * ri + q.y * rr;
                }
                a.coef[0][(size_t)r * kMaxKrylov + lane] = make_double2(rr, ri);
            }
            if (lane < m1) a.coef[1][(size_t)r * kMaxKrylov + lane] = make_double2(qv[lane].x, -qv[lane].y);
            if (lane == 0) a.w[r] = 1.0;
        }
        r++;
        __syncthreads();
    }
    if (tid == 0) a.out->n_nodes = r;
}

void launch_adjoint_coef(const AdjCoefArgs& a, double rank_tol, hipStream_t s) {
    hipLaunchKernelGGL(adjoint_coef_kernel, dim3(1), dim3(256), 0, s, a);
    if (rank_tol > 0.0) {
        if (std::max(a.m[0], a.m[1]) <= 40) hipLaunchKernelGGL(adjoint_rank_kernel<40>, dim3(1), dim3(256), 0, s, a, rank_tol);
        else hipLaunchKernelGGL(adjoint_rank_kernel<kMaxKrylov>, dim3(1), dim3(256), 0, s, a, rank_tol);
    }
}

// ---------------------------------------------------------------------------------------
// V w, V y_s and diagonal observables in one pass over the Krylov basis. All blocks sweep one
// front (tiles of kBlock rows dealt round-robin), and the basis columns are loaded
// non-temporally: m column streams read this way run at 5.5 TB/s on the MI355X against 4.9-5.1
// for a contiguous row range per block with ordinary loads.
// ---------------------------------------------------------------------------------------

// basis columns a thread loads before it uses any: 8, or 4 with 8 samples a launch (registers)
template <int SC>
constexpr int combine_burst() { return SC >= 8 ? 4 : 8; }

template <int SC, int NOB>
__global__ __launch_bounds__(kBlock) void combine_kernel(CombineArgs a) {
    __shared__ double2 cf[(1 + SC) * kMaxKrylov];
    __shared__ double2 sh[kBlock / 64];
    __shared__ double tabs[NOB > 0 ? NOB * kDiagTab : 1];   // value tables of byte-indexed diagonals
    const int tid = threadIdx.x;
    const int m = a.m, S = a.S;
    if (NOB > 0 && !a.diag && a.diag_tab)
        for (int i = tid; i < a.nobs * kDiagTab; i += kBlock) tabs[i] = a.diag_tab[i];
    for (int i = tid; i < (1 + SC) * kMaxKrylov; i += kBlock) {
        const int c = i / kMaxKrylov, j = i % kMaxKrylov;
        double2 v = make_double2(0.0, 0.0);
        if (j < m) {
            if (c == 0) {
                if (a.write_w) v = a.wcoef[j];
            } else if (c - 1 < S) {
                v = a.Y[(size_t)(a.y_base + c - 1) * kMaxKrylov + j];
            }
        }
        cf[i] = v;
    }
    __syncthreads();
    double acc[NOB > 0 ? NOB * SC : 1];
#pragma unroll
    for (int i = 0; i < (NOB > 0 ? NOB * SC : 1); i++) acc[i] = 0.0;
    for (int64_t row = (int64_t)blockIdx.x * kBlock + tid; row < a.n; row += (int64_t)gridDim.x * kBlock) {
        double2 aw = make_double2(0.0, 0.0);
        double2 as[SC];
#pragma unroll
        for (int s = 0; s < SC; s++) as[s] = make_double2(0.0, 0.0);
        auto add = [&](int j, double2 u) {
            if (a.write_w) {
                const double2 t = cmul(u, cf[j]);
                aw.x += t.x;
                aw.y += t.y;
            }
#pragma unroll
            for (int s = 0; s < SC; s++) {
                const double2 t = cmul(u, cf[(1 + s) * kMaxKrylov + j]);
                as[s].x += t.x;
                as[s].y += t.y;
            }
        };
        // a burst of basis columns in flight per thread: all loads of a burst are issued before
        // any is used (the sums still run over j in order)
        constexpr int B = combine_burst<SC>();
        int j = 0;
        for (; j + B <= m; j += B) {
            double2 u[B];
#pragma unroll
            for (int q = 0; q < B; q++) u[q] = load_stream(&a.U[(size_t)(j + q) * a.ld + row]);
#pragma unroll
            for (int q = 0; q < B; q++) add(j + q, u[q]);
        }
        for (; j < m; j++) add(j, load_stream(&a.U[(size_t)j * a.ld + row]));
        if (a.write_w) a.U0[row] = aw;
        if (a.write_last && S > 0) {
            double2 v = as[0];
#pragma unroll
            for (int s = 1; s < SC; s++)
                if (s == S - 1) v = as[s];
            a.last_out[row] = v;
        }
        if (NOB > 0) {
            double dg[NOB > 0 ? NOB : 1];                 // each observable's diagonal read once a row
            if (a.diag) {
#pragma unroll
                for (int o = 0; o < (NOB > 0 ? NOB : 1); o++) dg[o] = o < a.nobs ? a.diag[(size_t)o * a.n + row] : 0.0;
            } else {                                      // one-byte indices into the value tables
                typedef unsigned u4v __attribute__((ext_vector_type(4)));
                constexpr int NW = (NOB + 15) / 16;       // 16-byte words of indices per row
                u4v wv[NW > 0 ? NW : 1];
#pragma unroll
                for (int q = 0; q < NW; q++)
                    wv[q] = __builtin_nontemporal_load(
                        reinterpret_cast<const u4v*>(a.diag_idx + (size_t)row * a.idx_stride) + q);
#pragma unroll
                for (int o = 0; o < (NOB > 0 ? NOB : 1); o++) {
                    const unsigned word = wv[o / 16][(o % 16) / 4];
                    dg[o] = o < a.nobs ? tabs[o * kDiagTab + ((word >> (8 * (o % 4))) & 0xffu)] : 0.0;
                }
            }
#pragma unroll
            for (int s = 0; s < SC; s++) {
                const double pr = as[s].x * as[s].x + as[s].y * as[s].y;
#pragma unroll
                for (int o = 0; o < (NOB > 0 ? NOB : 1); o++)
                    if (o < a.nobs) acc[o * SC + s] += dg[o] * pr;
            }
        }
    }
    if (NOB == 0) return;
#pragma unroll
    for (int i = 0; i < (NOB > 0 ? NOB * SC : 1); i++) {
        const double2 t = block_sum2<kBlock>(make_double2(acc[i], 0.0), sh);
        if (tid == 0) a.partials[(size_t)blockIdx.x * (NOB * SC) + i] = t;
    }
}

// One block after combine_kernel<SC, NOB>: each observable's value at each sample, the nparts
// blocks' partials summed in block order (a launch of its own rather than a last-block
// reduction, which would cost every block of the combine an agent-scope release: an L2
// write-back on this multi-XCD part).
template <int SC, int NOB>
__global__ __launch_bounds__(kBlock) void combine_finish_kernel(CombineArgs a, int nparts) {
    __shared__ double2 sh[kBlock / 64];
    for (int i = 0; i < NOB * SC; i++) {
        const int o = i / SC, s = i % SC;
        const double2 tot = reduce_partials<kBlock>(a.partials + i, nparts, NOB * SC, sh);
        if (threadIdx.x == 0 && o < a.nobs && s < a.S)
            a.obs_values[(size_t)(a.obs_base + o) * a.n_samples + a.y_base + s] = tot.x;
    }
}

// ---------------------------------------------------------------------------------------
// The same pass for a step with more samples than combine_kernel<SC, NOB> takes at once, for
// up to kTiledObs diagonal observables: kTiledSC samples a pass, the observable sums through LDS
// instead of per-thread accumulators (which limit combine_kernel to NOB * SC <= 32). A block's
// rows go in tiles of kBlock, one row a thread as in combine_kernel (the same U w, U y_s and last
// sampled state, column by column in order); per tile each thread puts its row's |psi_s(r)|^2 and
// its observables' diagonal entries (IDX: their one-byte indices into the value tables; else the
// 8-byte values) in LDS, then thread t adds, for observable t % kTiledObs and every sample, the
// tile's rows of part t / kTiledObs in row order. After the last tile the parts of each
// (observable, sample) pair are added in part order into the block's partial;
// combine_tiled_finish_kernel adds the blocks' partials in block order. The two storage forms
// add the same doubles in the same order.
// ---------------------------------------------------------------------------------------
constexpr int kTiledSC = 8;
constexpr int kTiledObs = 16;
constexpr int kTiledParts = kBlock / kTiledObs;          // row parts of a tile (16 rows each)
constexpr int kTiledPairs = kTiledObs * kTiledSC;
constexpr int kTabPad = kDiagTab + 1;                   // a table row per observable, padded for LDS banks

template <bool IDX>
__global__ __launch_bounds__(kBlock) void combine_tiled_kernel(CombineArgs a) {
    __shared__ double2 cf[(1 + kTiledSC) * kMaxKrylov];
    __shared__ double pt[kBlock * (kTiledSC + 1)];       // |psi_s(r)|^2, row-major, padded
    __shared__ __align__(16) unsigned char ix[IDX ? kBlock * kTiledObs : 16];   // IDX: the rows' index bytes
    __shared__ double tabs[IDX ? kTiledObs * kTabPad : kBlock * (kTiledObs + 1)];   // tables, or the rows' values
    const int tid = threadIdx.x;
    const int m = a.m, S = a.S;
    if (IDX)
        for (int i = tid; i < kTiledObs * kTabPad; i += kBlock) {
            const int o = i / kTabPad, v = i % kTabPad;
            tabs[i] = (o < a.nobs && v < kDiagTab) ? a.diag_tab[o * kDiagTab + v] : 0.0;
        }
    for (int i = tid; i < (1 + kTiledSC) * kMaxKrylov; i += kBlock) {
        const int c = i / kMaxKrylov, j = i % kMaxKrylov;
        double2 v = make_double2(0.0, 0.0);
        if (j < m) {
            if (c == 0) {
                if (a.write_w) v = a.wcoef[j];
            } else if (c - 1 < S) {
                v = a.Y[(size_t)(a.y_base + c - 1) * kMaxKrylov + j];
            }
        }
        cf[i] = v;
    }
    const int ro = tid % kTiledObs, rp = tid / kTiledObs;   // this thread's observable and row part
    double acc[kTiledSC];
#pragma unroll
    for (int s = 0; s < kTiledSC; s++) acc[s] = 0.0;
    __syncthreads();
    for (int64_t t0 = (int64_t)blockIdx.x * kBlock; t0 < a.n; t0 += (int64_t)gridDim.x * kBlock) {
        const int64_t row = t0 + tid;
        double2 aw = make_double2(0.0, 0.0);
        double2 as[kTiledSC];
#pragma unroll
        for (int s = 0; s < kTiledSC; s++) as[s] = make_double2(0.0, 0.0);
        typedef unsigned u4v __attribute__((ext_vector_type(4)));
        u4v idx = {0u, 0u, 0u, 0u};
        if (row < a.n) {
            auto add = [&](int j, double2 u) {
                if (a.write_w) {
                    const double2 t = cmul(u, cf[j]);
                    aw.x += t.x;
                    aw.y += t.y;
                }
#pragma unroll
                for (int s = 0; s < kTiledSC; s++) {
                    const double2 t = cmul(u, cf[(1 + s) * kMaxKrylov + j]);
                    as[s].x += t.x;
                    as[s].y += t.y;
                }
            };
            constexpr int B = combine_burst<kTiledSC>();
            int j = 0;
            for (; j + B <= m; j += B) {
                double2 u[B];
#pragma unroll
                for (int q = 0; q < B; q++) u[q] = load_stream(&a.U[(size_t)(j + q) * a.ld + row]);
#pragma unroll
                for (int q = 0; q < B; q++) add(j + q, u[q]);
            }
            for (; j < m; j++) add(j, load_stream(&a.U[(size_t)j * a.ld + row]));
            if (a.write_w) a.U0[row] = aw;
            if (a.write_last && S > 0) {
                double2 v = as[0];
#pragma unroll
                for (int s = 1; s < kTiledSC; s++)
                    if (s == S - 1) v = as[s];
                a.last_out[row] = v;
            }
            if (IDX)
                idx = __builtin_nontemporal_load(
                    reinterpret_cast<const u4v*>(a.diag_idx + (size_t)row * a.idx_stride));
        }
#pragma unroll
        for (int s = 0; s < kTiledSC; s++) pt[tid * (kTiledSC + 1) + s] = as[s].x * as[s].x + as[s].y * as[s].y;
        if (IDX) {
            *reinterpret_cast<u4v*>(ix + tid * kTiledObs) = idx;
        } else {
#pragma unroll
            for (int o = 0; o < kTiledObs; o++)
                tabs[tid * (kTiledObs + 1) + o] = (row < a.n && o < a.nobs) ? a.diag[(size_t)o * a.n + row] : 0.0;
        }
        __syncthreads();
        for (int k = 0; k < kBlock / kTiledParts; k++) {
            const int i = rp * (kBlock / kTiledParts) + k;
            const double d = IDX ? tabs[ro * kTabPad + ix[i * kTiledObs + ro]] : tabs[i * (kTiledObs + 1) + ro];
#pragma unroll
            for (int s = 0; s < kTiledSC; s++) acc[s] += d * pt[i * (kTiledSC + 1) + s];
        }
        __syncthreads();                                  // pt and ix are refilled by the next tile
    }
    // parts of each pair in part order (pt is free again: kTiledParts x kTiledPairs doubles fit)
    static_assert(kTiledParts * kTiledPairs <= kBlock * (kTiledSC + 1), "the parts fit in pt");
#pragma unroll
    for (int s = 0; s < kTiledSC; s++) pt[rp * kTiledPairs + ro * kTiledSC + s] = acc[s];
    __syncthreads();
    if (tid < kTiledPairs) {
        double t = 0.0;
        for (int q = 0; q < kTiledParts; q++) t += pt[q * kTiledPairs + tid];
        a.partials[(size_t)blockIdx.x * kTiledPairs + tid] = make_double2(t, 0.0);
    }
}

// One block after combine_tiled_kernel: thread p adds pair p's block partials in block order.
__global__ __launch_bounds__(kBlock) void combine_tiled_finish_kernel(CombineArgs a, int nparts) {
    const int p = threadIdx.x;
    if (p >= kTiledPairs) return;
    double t = 0.0;
    for (int b = 0; b < nparts; b++) t += a.partials[(size_t)b * kTiledPairs + p].x;
    const int o = p / kTiledSC, s = p % kTiledSC;
    if (o < a.nobs && s < a.S) a.obs_values[(size_t)(a.obs_base + o) * a.n_samples + a.y_base + s] = t;
}

__global__ void store_obs_kernel(const double2* res, int slot, int mode, double* out) {
    const double2 r = res[slot];
    *out = mode == 0 ? r.x : r.x * r.x + r.y * r.y;
}

// ---------------------------------------------------------------------------------------
void launch_step_control(const StepConsts& c, StepState* st, KrylovScalars* ks, hipStream_t s) {
    hipLaunchKernelGGL(step_control_kernel, dim3(1), dim3(kCtl), 0, s, c, st, ks);
}

void launch_batch_step_control(const StepConsts* c, StepState* st, KrylovScalars* ks, BatchCtl* bc, int n_inst,
                               long long n_samples, hipStream_t s) {
    hipLaunchKernelGGL(batch_step_control_kernel, dim3(n_inst), dim3(kCtl), 0, s, c, st, ks, bc, n_samples);
}

int combine_obs_per_launch() { return 32; }

int combine_samples_per_launch(int nobs) {
    if (nobs <= 4) return 8;
    if (nobs <= 8) return 4;
    if (nobs <= 16) return 2;
    return 1;
}

template <int SC, int NOB>
static void launch_combine_t(const CombineArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((combine_kernel<SC, NOB>), dim3(grid), dim3(kBlock), 0, s, a);
    if constexpr (NOB > 0)
        hipLaunchKernelGGL((combine_finish_kernel<SC, NOB>), dim3(1), dim3(kBlock), 0, s, a, grid);
}

void launch_combine(const CombineArgs& a, int grid, hipStream_t s) {
    if (a.nobs == 0) launch_combine_t<8, 0>(a, grid, s);
    else if (a.nobs <= 1) launch_combine_t<8, 1>(a, grid, s);
    else if (a.nobs <= 2) launch_combine_t<8, 2>(a, grid, s);
    else if (a.nobs <= 4) launch_combine_t<8, 4>(a, grid, s);
    else if (a.nobs <= 8) launch_combine_t<4, 8>(a, grid, s);
    else if (a.nobs <= 16) launch_combine_t<2, 16>(a, grid, s);
    else launch_combine_t<1, 32>(a, grid, s);
}

int combine_tiled_samples(int nobs) { return nobs >= 1 && nobs <= kTiledObs ? kTiledSC : 0; }

void launch_combine_tiled(const CombineArgs& a, int grid, hipStream_t s) {
    if (a.diag) hipLaunchKernelGGL(combine_tiled_kernel<false>, dim3(grid), dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL(combine_tiled_kernel<true>, dim3(grid), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(combine_tiled_finish_kernel, dim3(1), dim3(kBlock), 0, s, a, grid);
}

void launch_store_obs(const double2* res, int slot, int mode, double* out, hipStream_t s) {
    hipLaunchKernelGGL(store_obs_kernel, dim3(1), dim3(1), 0, s, res, slot, mode, out);
}

}  // namespace te
