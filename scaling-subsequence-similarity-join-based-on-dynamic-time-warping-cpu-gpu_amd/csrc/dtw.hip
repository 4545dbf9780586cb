// Banded DTW over an explicit pair list, one wavefront per pair.
//  * prec 32: FP32 cells on z = (ac - mu_c) * isig (the join's own arithmetic);
//  * prec 64: FP64 cells on z = (a - mu) / sigma (the re-verification of the join's winners;
//    same formula as the reference's z-normalised subsequences,
//    CPU_version1/new_dtw_motifGUI.cpp:41-44).
#include "dtw_device.hpp"
#include "../../include/dtwj.h"
#include "internal.hpp"

namespace dtwj {

namespace {

template <typename T, int QP>
__global__ __launch_bounds__(64) void k_dtw_pairs(JoinArgs A, const double *__restrict__ sig64,
                                                  const int64_t *__restrict__ pi, const int64_t *__restrict__ pj,
                                                  int64_t npairs, double *cost) {
    extern __shared__ unsigned char smem_raw[];
    T *zi = reinterpret_cast<T *>(smem_raw);
    T *zj = zi + A.m;
    const int64_t p = blockIdx.x;
    if (p >= npairs) return;
    const int64_t i = pi[p], j = pj[p];
    const int lane = threadIdx.x;
    if (sizeof(T) == 8) {
        const double mi = A.mu[i], si = sig64[i], mj = A.mu[j], sj = sig64[j];
        for (int k = lane; k < A.m; k += 64) {
            zi[k] = (T)((A.a[i + k] - mi) / si);
            zj[k] = (T)((A.a[j + k] - mj) / sj);
        }
    } else {
        const float mi = A.mu_c[i], ii = A.win[i].isig, mj = A.mu_c[j], ij = A.win[j].isig;
        for (int k = lane; k < A.m; k += 64) {
            zi[k] = (T)((A.ac[i + k] - mi) * ii);
            zj[k] = (T)((A.ac[j + k] - mj) * ij);
        }
    }
    __syncthreads();
    const PlainZ<T> z{zi, zj};
    int t_stop = 0;
    const T r = dtw_band<T, QP>(A.m, A.w, 64, 1, dev_inf<T>(), z, t_stop);
    if (lane == 0) cost[p] = (double)r;
}

template <typename T>
int launch_pairs_t(const JoinArgs &A, const double *sig64, const int64_t *pi, const int64_t *pj, int64_t npairs,
                   double *cost, hipStream_t s) {
    const int QP = dtw_qp_for(A.w);
    const size_t lds = 2 * (size_t)A.m * sizeof(T);
    const dim3 grid((unsigned)npairs), block(64);
    switch (QP) {
        case 1: hipLaunchKernelGGL((k_dtw_pairs<T, 1>), grid, block, lds, s, A, sig64, pi, pj, npairs, cost); break;
        case 2: hipLaunchKernelGGL((k_dtw_pairs<T, 2>), grid, block, lds, s, A, sig64, pi, pj, npairs, cost); break;
        case 4: hipLaunchKernelGGL((k_dtw_pairs<T, 4>), grid, block, lds, s, A, sig64, pi, pj, npairs, cost); break;
        case 8: hipLaunchKernelGGL((k_dtw_pairs<T, 8>), grid, block, lds, s, A, sig64, pi, pj, npairs, cost); break;
        default: set_error("band too wide: 2w+1 must be <= 1024"); return DTWJ_EINVAL;
    }
    DTWJ_HIP(hipGetLastError());
    return 0;
}

// ---- FP64 re-scoring of a join's candidate list.  One wave per candidate: candidates whose
// FP32 cost exceeds ub (only a near-tie of the best can be the FP64 winner) get +inf, the
// others their FP64 cost.  Duplicates (a pair pushed by the seed and by the join) need no
// removal: the minimum below is unchanged by them.
template <int QP>
__global__ __launch_bounds__(64) void k_rescore(JoinArgs A, const double *__restrict__ sig64,
                                                const int4 *__restrict__ cands, int64_t n, float ub,
                                                double *cost64) {
    extern __shared__ unsigned char smem_raw[];
    double *zi = reinterpret_cast<double *>(smem_raw);
    double *zj = zi + A.m;
    const int64_t p = blockIdx.x;
    if (p >= n) return;
    const int4 c = cands[p];
    const int lane = threadIdx.x;
    if (!(__int_as_float(c.z) <= ub)) {
        if (lane == 0) cost64[p] = __builtin_huge_val();
        return;
    }
    const int64_t i = c.x, j = c.y;
    const double mi = A.mu[i], si = sig64[i], mj = A.mu[j], sj = sig64[j];
    for (int k = lane; k < A.m; k += 64) {
        zi[k] = (A.a[i + k] - mi) / si;
        zj[k] = (A.a[j + k] - mj) / sj;
    }
    __syncthreads();
    const PlainZ<double> z{zi, zj};
    int t_stop = 0;
    const double r = dtw_band<double, QP>(A.m, A.w, 64, 1, dev_inf<double>(), z, t_stop);
    if (lane == 0) cost64[p] = r;
}

// (cost, i, j, FP32 cost) in lexicographic order: the FP64 winner with the smallest i, then j.
__device__ inline bool winner_less(const Winner &a, const Winner &b) {
    if (a.cost != b.cost) return a.cost < b.cost;
    if (a.i != b.i) return a.i < b.i;
    if (a.j != b.j) return a.j < b.j;
    return a.cost32 < b.cost32;
}

// Per-block minimum over a strided slice of the re-scored candidates (and the count of
// candidates that were re-scored), then one block over the partials.
constexpr int kArgminThreads = 256;
constexpr int kArgminBlocks = 1024;

__device__ inline void block_min(Winner &w) {
    __shared__ Winner sh[kArgminThreads];
    const int t = threadIdx.x;
    sh[t] = w;
    __syncthreads();
    for (int off = kArgminThreads / 2; off > 0; off >>= 1) {
        if (t < off) {
            Winner o = sh[t + off];
            Winner &me = sh[t];
            const unsigned long long nv = me.verified + o.verified;
            if (winner_less(o, me)) me = o;
            me.verified = nv;
        }
        __syncthreads();
    }
    w = sh[0];
}

__global__ __launch_bounds__(kArgminThreads) void k_argmin_part(const int4 *__restrict__ cands,
                                                                const double *__restrict__ cost64, int64_t n,
                                                                Winner *part) {
    Winner w{__builtin_huge_val(), 0x7fffffff, 0x7fffffff, __builtin_huge_valf(), 0u, 0ull};
    for (int64_t p = (int64_t)blockIdx.x * kArgminThreads + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * kArgminThreads) {
        const double c = cost64[p];
        if (c < __builtin_huge_val()) {
            const int4 e = cands[p];
            const Winner x{c, e.x, e.y, __int_as_float(e.z), 0u, 0ull};
            if (winner_less(x, w)) {
                const unsigned long long nv = w.verified;
                w = x;
                w.verified = nv;
            }
            ++w.verified;
        }
    }
    block_min(w);
    if (threadIdx.x == 0) part[blockIdx.x] = w;
}

__global__ __launch_bounds__(kArgminThreads) void k_argmin_final(const Winner *__restrict__ part, int nparts,
                                                                 Winner *out) {
    Winner w{__builtin_huge_val(), 0x7fffffff, 0x7fffffff, __builtin_huge_valf(), 0u, 0ull};
    for (int p = threadIdx.x; p < nparts; p += kArgminThreads) {
        const Winner x = part[p];
        const unsigned long long nv = w.verified + x.verified;
        if (winner_less(x, w)) w = x;
        w.verified = nv;
    }
    block_min(w);
    if (threadIdx.x == 0) *out = w;
}

}  // namespace

int rescore_partials() { return kArgminBlocks; }

int launch_rescore(const JoinArgs &A, const double *sig64, const int4 *cands, int64_t n, float ub, double *cost64,
                   Winner *part, Winner *out, hipStream_t s) {
    if (n <= 0) return 0;
    const int QP = dtw_qp_for(A.w);
    const size_t lds = 2 * (size_t)A.m * sizeof(double);
    // one block per candidate: the grid's x dimension holds up to 2^31 - 1 blocks
    if (n > 0x7fffffffLL) {
        set_error("more than 2^31 - 1 candidates to re-score");
        return DTWJ_EOVERFLOW;
    }
    const dim3 grid((unsigned)n), block(64);
    switch (QP) {
        case 1: hipLaunchKernelGGL(k_rescore<1>, grid, block, lds, s, A, sig64, cands, n, ub, cost64); break;
        case 2: hipLaunchKernelGGL(k_rescore<2>, grid, block, lds, s, A, sig64, cands, n, ub, cost64); break;
        case 4: hipLaunchKernelGGL(k_rescore<4>, grid, block, lds, s, A, sig64, cands, n, ub, cost64); break;
        case 8: hipLaunchKernelGGL(k_rescore<8>, grid, block, lds, s, A, sig64, cands, n, ub, cost64); break;
        default: set_error("band too wide: 2w+1 must be <= 1024"); return DTWJ_EINVAL;
    }
    DTWJ_HIP(hipGetLastError());
    const int64_t want = (n + kArgminThreads - 1) / kArgminThreads;
    const int nb = (int)(want < kArgminBlocks ? want : kArgminBlocks);
    hipLaunchKernelGGL(k_argmin_part, dim3(nb), dim3(kArgminThreads), 0, s, cands, cost64, n, part);
    DTWJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_argmin_final, dim3(1), dim3(kArgminThreads), 0, s, part, nb, out);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_dtw_pairs(const JoinArgs &A, const double *sig64, const int64_t *pi, const int64_t *pj, int64_t npairs,
                     int32_t prec, double *cost, hipStream_t s) {
    if (npairs <= 0) return 0;
    if (prec == 64) return launch_pairs_t<double>(A, sig64, pi, pj, npairs, cost, s);
    return launch_pairs_t<float>(A, sig64, pi, pj, npairs, cost, s);
}

// LB_Keogh of window j against the radius-w envelope of window i (all m positions) plus the
// LB_Petitjean increments (CPU_version1/LB_PETITJEAN.cpp:9-123; q = window j, x = window i):
// project q into x's envelope on positions [3, m - 3), take the radius-w envelope of the
// projection, and add, per position of [3, m - 3), what x pays outside it beyond what q's own
// envelope already accounts for.  FP64, one block of 256 threads per pair, z-windows in LDS;
// every envelope value is a direct max / min over its clipped radius-w range.
__global__ __launch_bounds__(256) void k_lb_petitjean(const double *__restrict__ a, const double *__restrict__ mu,
                                                      const double *__restrict__ sig, int m, int w,
                                                      const int64_t *__restrict__ pi, const int64_t *__restrict__ pj,
                                                      double *__restrict__ out) {
    extern __shared__ double sh[];
    double *zi = sh, *zj = sh + m, *pr = sh + 2 * m;
    __shared__ double red[256];
    const int t = threadIdx.x;
    const int64_t i = pi[blockIdx.x], j = pj[blockIdx.x];
    const double mi = mu[i], si = sig[i], mj = mu[j], sj = sig[j];
    for (int k = t; k < m; k += 256) {
        zi[k] = (a[i + k] - mi) / si;
        zj[k] = (a[j + k] - mj) / sj;
    }
    __syncthreads();
    auto env = [&](const double *z, int k, double &u, double &l) {
        const int lo = k - w > 0 ? k - w : 0, hi = k + w < m - 1 ? k + w : m - 1;
        u = l = z[lo];
        for (int r = lo + 1; r <= hi; ++r) {
            u = fmax(u, z[r]);
            l = fmin(l, z[r]);
        }
    };
    auto sq = [](double x) { return x * x; };
    double acc = 0.0;
    for (int k = t; k < m; k += 256) {
        double u, l;
        env(zi, k, u, l);
        const double q = zj[k];
        acc += q > u ? sq(q - u) : (q < l ? sq(q - l) : 0.0);
        double p = q;
        if (k >= 3 && k < m - 3) p = q > u ? u : (q < l ? l : q);
        pr[k] = p;
    }
    __syncthreads();
    for (int k = t + 3; k < m - 3; k += 256) {
        double up, lp, uq, lq;
        env(pr, k, up, lp);
        env(zj, k, uq, lq);
        const double x = zi[k];
        double d = 0.0;
        if (x > up && up > uq)
            d = sq(x - uq) - sq(up - uq);
        else if (x < lp && lp < lq)
            d = sq(x - lq) - sq(lp - lq);
        else if (x > up)
            d = sq(x - up);
        else if (x < lp)
            d = sq(x - lp);
        acc += d;
    }
    red[t] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    if (t == 0) out[blockIdx.x] = red[0];
}

int launch_lb_petitjean(const double *a, const double *mu, const double *sig, int32_t m, int32_t w, const int64_t *pi,
                        const int64_t *pj, int64_t npairs, double *out, hipStream_t s) {
    if (npairs <= 0) return 0;
    if (npairs > 0x7fffffffLL || m > 4096) {
        set_error("LB_Petitjean pairs: at most 2^31 - 1 pairs and m <= 4096");
        return DTWJ_EINVAL;
    }
    hipLaunchKernelGGL(k_lb_petitjean, dim3((unsigned)npairs), dim3(256), sizeof(double) * 3 * (size_t)m, s, a, mu, sig,
                       m, w, pi, pj, out);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

}  // namespace dtwj
