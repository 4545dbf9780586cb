// The reference's global occupancy mask and its masked ball bound on gfx950.
//
// Mask (CPU_version1/compute_shared_data_global.cpp:12-46, CPU_version1/compute_MASK_global.cpp:11-93,
// GPU_version1/src/compute_MASK_global.cu:92-171, table length as
// CPU_version1/new_dtw_motifGUI.cpp:151-176): for every window offset k a histogram of the
// z-values z_i[k] of all windows i in bins of 0.01 (bin = floor(100 z + len / 2), clamped to
// [1, len], len = max(600, 200 floor(max |z|))), turned into a cumulative count; for every
// (window i, offset k) the number of windows whose value at offset k lies inside that window's
// own envelope there, [Lhat_i[k], Uhat_i[k]] (radius w, clipped at the window's ends), divided by
// the number of windows and summed at series position i + k; divided by the coverage of the
// position (1..m, m, ..m..1); MASK = 1 where that is at most its mean over the series.
//
// Every step follows the reference's FP64 operations in its order, so that the mask is the one
// the reference computes from the same window statistics: the z-values as (x - mu) / sigma, the
// bin expression without contraction, each position's sum over its windows in increasing window
// order, and the mean as one sequential sum.  Kernels:
//   k_mask_zmax   largest |z| (the extremes of each window, O(m) per window)
//   k_mask_len    table length
//   k_mask_hist   per offset: LDS histogram of a slice of the windows, merged with atomics
//   k_mask_cdf    per offset: cumulative counts
//   k_mask_sum    per position: the window sum, the clipped envelope of a window at an edge
//                 offset from a sparse table (range max / min) of the block's span in LDS
//   k_mask_mean   sequential mean (one wave), then the mask
//
// Masked ball bound (CPU_version1/diag_mask_global.cpp:11-226): for the envelope window e and the
// data window d, over K = {k in [3, m-4] : MASK[e + k] = 1},
//   dist = sum_K (z_d[k] - Uhat_e[k])^2 + (z_d[k] - Lhat_e[k])^2,  DUL2 = sum over the masked
//   positions of [0, m-1] of ((U - L)[e + k] / sigma_e)^2,
//   lb = 0 if dist < DUL2, else (sqrt(2 dist - DUL2) - sqrt(DUL2)) / 2,
// with the raw-series envelope U, L (a superset of the window's own: still a lower bound of the
// banded DTW).  k_masked_ball evaluates it for explicit pairs by direct FP64 sums (the reference
// keeps running sums along a diagonal), one wave per pair, both directions.
#include <float.h>

#include "internal.hpp"

namespace dtwj {

namespace {

struct MaskArgs {
    const double *a, *U, *L, *mu, *sig;
    int64_t n, nw;
    int32_t m, w;
    int32_t lmax;  // bins allocated per offset
};

__device__ inline unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

// Largest |z| over every window and offset: max over windows of |z(max)|, |z(min)|, which is
// the largest |z| itself because (x - mu) / sigma is monotone in x under rounding.
__global__ __launch_bounds__(256) void k_mask_zmax(MaskArgs A, unsigned long long *zmax_bits) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double zm = 0.0;
    if (i < A.nw) {
        double hi = -DBL_MAX, lo = DBL_MAX;
        const double *x = A.a + i;
        for (int k = 0; k < A.m; ++k) {
            const double v = x[k];
            hi = fmax(hi, v);
            lo = fmin(lo, v);
        }
        const double mu = A.mu[i], sg = A.sig[i];
        zm = fmax(fabs((hi - mu) / sg), fabs((lo - mu) / sg));
    }
    for (int off = 32; off > 0; off >>= 1) zm = fmax(zm, __shfl_xor(zm, off, 64));
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = zm;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
        atomicMax(zmax_bits, dbits(b));  // non-negative doubles order like their bits
    }
}

__global__ void k_mask_init(unsigned long long *zmax_bits, int32_t *status) {
    *zmax_bits = 0ull;
    status[0] = 0;  // table length
    status[1] = 0;  // error flag
}

// len = max(600, floor(max |z|) * 2 * 100) (CPU_version1/new_dtw_motifGUI.cpp:173-176)
__global__ void k_mask_len(const unsigned long long *zmax_bits, int32_t lmax, int32_t *status) {
    const double zmax = __longlong_as_double((long long)*zmax_bits);
    const double v = floor(zmax) * 2 * 100;
    const double len = v > 600.0 ? v : 600.0;
    if (!(len <= (double)lmax)) {
        status[1] = 1;
        status[0] = 600;
    } else {
        status[0] = (int32_t)len;
    }
}

// bin of a value, clamped to [1, len] (FIND_POS_GLOBAL, CPU_version1/alldef/matrix.h:17)
__device__ inline int find_pos(double z, int len) {
    DTWJ_NO_CONTRACT
    const double p = floor(z * 100 + len * 0.5);
    if (p <= 0) return 1;
    if (p > len) return len;
    return (int)p;
}

// Histogram of z_i[k] over the windows [i0, i1) of slice blockIdx.x, k = blockIdx.y.
__global__ __launch_bounds__(256) void k_mask_hist(MaskArgs A, const int32_t *status, int64_t slice,
                                                   uint32_t *table) {
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    uint32_t *h = reinterpret_cast<uint32_t *>(dyn_lds);
    const int len = status[0];
    const int k = (int)blockIdx.y;
    for (int b = threadIdx.x; b < len; b += 256) h[b] = 0u;
    __syncthreads();
    const int64_t i0 = (int64_t)blockIdx.x * slice;
    const int64_t i1 = i0 + slice < A.nw ? i0 + slice : A.nw;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
        const double z = (A.a[i + k] - A.mu[i]) / A.sig[i];
        atomicAdd(&h[find_pos(z, len) - 1], 1u);
    }
    __syncthreads();
    uint32_t *row = table + (size_t)k * A.lmax;
    for (int b = threadIdx.x; b < len; b += 256)
        if (h[b]) atomicAdd(&row[b], h[b]);
}

// Cumulative counts of each offset's histogram (one block per offset).
__global__ __launch_bounds__(256) void k_mask_cdf(MaskArgs A, const int32_t *status, uint32_t *table) {
    __shared__ uint32_t part[256];
    const int len = status[0];
    uint32_t *row = table + (size_t)blockIdx.x * A.lmax;
    const int per = (len + 255) / 256;
    const int b0 = threadIdx.x * per, b1 = b0 + per < len ? b0 + per : len;
    uint32_t s = 0;
    for (int b = b0; b < b1; ++b) s += row[b];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {  // inclusive scan of the segment sums
        const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t c = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0u;
    for (int b = b0; b < b1; ++b) {
        c += row[b];
        row[b] = c;
    }
}

constexpr int kSumBlock = 256;

// Sparse-table levels for ranges up to qmax positions.
__host__ __device__ inline int mask_levels(int qmax) {
    int p = 0;
    while ((2 << p) <= qmax) ++p;
    return p + 1;
}

// Occupancy statistic of every position t of the block: sum over the windows i = t - k (k from
// the largest offset down, i.e. windows in increasing order) of (cdf[k][pos(Uhat)] -
// cdf[k][pos(Lhat)]) / nw, divided by the coverage of t.
__global__ __launch_bounds__(kSumBlock) void k_mask_sum(MaskArgs A, const int32_t *status, const uint32_t *table,
                                                        int levels, double *diff) {
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    double *st = reinterpret_cast<double *>(dyn_lds);
    const int len = status[0];
    const int w = A.w, m = A.m;
    const int64_t n = A.n, nw = A.nw;
    const int64_t t0 = (int64_t)blockIdx.x * kSumBlock;
    const int64_t base = t0 - w;
    const int span = kSumBlock + 2 * w;
    double *mx = st, *mn = st + (size_t)levels * span;
    for (int x = threadIdx.x; x < span; x += kSumBlock) {
        const int64_t p = base + x;
        const bool in = p >= 0 && p < n;
        const double v = in ? A.a[p] : 0.0;
        mx[x] = in ? v : -DBL_MAX;
        mn[x] = in ? v : DBL_MAX;
    }
    __syncthreads();
    for (int lv = 1; lv < levels; ++lv) {
        const int h = 1 << (lv - 1);
        for (int x = threadIdx.x; x + 2 * h <= span; x += kSumBlock) {
            mx[lv * span + x] = fmax(mx[(lv - 1) * span + x], mx[(lv - 1) * span + x + h]);
            mn[lv * span + x] = fmin(mn[(lv - 1) * span + x], mn[(lv - 1) * span + x + h]);
        }
        __syncthreads();
    }
    const int64_t t = t0 + threadIdx.x;
    if (t >= n) return;
    const int k_hi = (int)(t < (int64_t)(m - 1) ? t : (int64_t)(m - 1));
    const int k_lo = (int)(t - (nw - 1) > 0 ? t - (nw - 1) : 0);
    const double Ut = A.U[t], Lt = A.L[t];
    const double dnw = (double)nw;
    double nd = 0.0;
    for (int k = k_hi; k >= k_lo; --k) {
        const int64_t i = t - k;
        const int lk = k < w ? k : w, rk = (m - 1 - k) < w ? (m - 1 - k) : w;
        double uv = Ut, lv = Lt;
        if (lk != w || rk != w) {
            // range [t - lk, t + rk] of the block's span (it lies inside the series)
            const int l = (int)(t - lk - base), r = (int)(t + rk - base);
            const int q = r - l + 1;
            const int p = 31 - __builtin_clz((unsigned)q);
            const int r2 = r - (1 << p) + 1;
            uv = fmax(mx[p * span + l], mx[p * span + r2]);
            lv = fmin(mn[p * span + l], mn[p * span + r2]);
        }
        const double mu = A.mu[i], sg = A.sig[i];
        const int pu = find_pos((uv - mu) / sg, len), pl = find_pos((lv - mu) / sg, len);
        const uint32_t *row = table + (size_t)k * A.lmax;
        const double d2 = (double)row[pu - 1] - (double)row[pl - 1];
        nd += d2 / dnw;
    }
    // coverage ramp [1..m, m x (n - 2m), m..1] (CPU_version1/compute_MASK_global.cpp:15-23)
    double mod;
    if (t < m)
        mod = (double)(t + 1);
    else if (t < n - m)
        mod = (double)m;
    else
        mod = (double)m - (double)(t - (n - m));
    diff[t] = nd / mod;
}

// Mean of the statistic as one sequential sum (CPU_version1/allmatrixop/Mean.cpp:5-18), on one
// wave: 64 values are loaded at a time and added in order from the lanes that hold them
// (v_readlane, no LDS round trip per value).
__device__ inline double lane_value(double v, int q) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, q), hi = __builtin_amdgcn_readlane((int)(b >> 32), q);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__global__ __launch_bounds__(64) void k_mask_mean(const double *diff, int64_t n, double *ths_out) {
    const int lane = threadIdx.x;
    double sum = 0.0;
    int64_t b = 0;
    for (; b + 64 <= n; b += 64) {
        const double v = diff[b + lane];
#pragma unroll
        for (int q = 0; q < 64; ++q) sum += lane_value(v, q);
    }
    const double v = b + lane < n ? diff[b + lane] : 0.0;
    for (int q = 0; q < (int)(n - b); ++q) sum += lane_value(v, q);
    if (lane == 0) *ths_out = sum / (double)n;
}

__global__ __launch_bounds__(256) void k_mask_final(const double *diff, int64_t n, const double *ths, uint8_t *mask) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) mask[t] = diff[t] <= *ths ? 1 : 0;
}

// Masked ball bound of explicit pairs, both directions: lb[2 p] (envelope of pi[p] against
// window pj[p]) and lb[2 p + 1] (the reverse).  One wave per pair.
__global__ __launch_bounds__(64) void k_masked_ball(MaskArgs A, const uint8_t *mask, const int64_t *pi,
                                                    const int64_t *pj, int64_t npairs, double *lb) {
    const int64_t p = blockIdx.x;
    if (p >= npairs) return;
    const int lane = threadIdx.x;
    const int m = A.m;
    for (int dir = 0; dir < 2; ++dir) {
        const int64_t e = dir == 0 ? pi[p] : pj[p], d = dir == 0 ? pj[p] : pi[p];
        const double mue = A.mu[e], ise = 1.0 / A.sig[e], mud = A.mu[d], isd = 1.0 / A.sig[d];
        double dist = 0.0, dul2 = 0.0;
        for (int k = lane; k < m; k += 64) {
            if (!mask[e + k]) continue;
            const double u = A.U[e + k], l = A.L[e + k];
            const double h = (u - l) * ise;
            dul2 += h * h;
            if (k >= 3 && k <= m - 4) {
                const double z = (A.a[d + k] - mud) * isd;
                const double du = z - (u - mue) * ise, dl = z - (l - mue) * ise;
                dist += du * du + dl * dl;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            dist += __shfl_xor(dist, off, 64);
            dul2 += __shfl_xor(dul2, off, 64);
        }
        if (lane == 0) lb[2 * p + dir] = dist < dul2 ? 0.0 : 0.5 * (sqrt(2 * dist - dul2) - sqrt(dul2));
    }
}

}  // namespace

int64_t mask_table_bins(int32_t m) {
    int64_t r = 0;
    while ((r + 1) * (r + 1) <= (int64_t)(m - 1)) ++r;  // floor(sqrt(m - 1)): |z| <= sqrt(m - 1)
    const int64_t b = 200 * (r + 1);                      // one more for the rounding of sigma
    return b > 600 ? b : 600;
}

size_t mask_sum_lds_bytes(int32_t m, int32_t w) {
    const int qmax = 2 * w < m ? 2 * w : m;
    const int levels = w > 0 ? mask_levels(qmax) : 1;
    return (size_t)2 * levels * (kSumBlock + 2 * w) * sizeof(double);
}

int launch_global_mask(const MaskBuffers &B, hipStream_t s) {
    MaskArgs A{B.a, B.U, B.L, B.mu, B.sig, B.n, B.n - B.m + 1, B.m, B.w, (int32_t)B.lmax};
    const int64_t nw = A.nw;
    hipLaunchKernelGGL(k_mask_init, dim3(1), dim3(1), 0, s, B.zmax_bits, B.status);
    hipLaunchKernelGGL(k_mask_zmax, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, A, B.zmax_bits);
    hipLaunchKernelGGL(k_mask_len, dim3(1), dim3(1), 0, s, B.zmax_bits, A.lmax, B.status);
    DTWJ_HIP(hipMemsetAsync(B.table, 0, sizeof(uint32_t) * (size_t)B.lmax * B.m, s));
    // about 4096 blocks: slices of the windows per offset
    int64_t slices = 4096 / B.m;
    if (slices < 1) slices = 1;
    const int64_t max_slices = (nw + 4095) / 4096;
    if (slices > max_slices) slices = max_slices;
    const int64_t slice = (nw + slices - 1) / slices;
    // dynamic LDS beyond 64 KB must be allowed per kernel (up to 160 KB on gfx950)
    DTWJ_HIP(hipFuncSetAttribute((const void *)k_mask_hist, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(sizeof(uint32_t) * B.lmax)));
    DTWJ_HIP(hipFuncSetAttribute((const void *)k_mask_sum, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)mask_sum_lds_bytes(B.m, B.w)));
    hipLaunchKernelGGL(k_mask_hist, dim3((unsigned)slices, (unsigned)B.m), dim3(256), sizeof(uint32_t) * B.lmax, s, A,
                       (const int32_t *)B.status, slice, B.table);
    hipLaunchKernelGGL(k_mask_cdf, dim3((unsigned)B.m), dim3(256), 0, s, A, (const int32_t *)B.status, B.table);
    const int qmax = 2 * B.w < B.m ? 2 * B.w : B.m;
    const int levels = B.w > 0 ? mask_levels(qmax) : 1;
    hipLaunchKernelGGL(k_mask_sum, dim3((unsigned)((B.n + kSumBlock - 1) / kSumBlock)), dim3(kSumBlock),
                       mask_sum_lds_bytes(B.m, B.w), s, A, (const int32_t *)B.status, (const uint32_t *)B.table,
                       levels, B.diff);
    hipLaunchKernelGGL(k_mask_mean, dim3(1), dim3(64), 0, s, (const double *)B.diff, B.n, B.ths);
    hipLaunchKernelGGL(k_mask_final, dim3((unsigned)((B.n + 255) / 256)), dim3(256), 0, s, (const double *)B.diff,
                       B.n, (const double *)B.ths, B.mask);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_masked_ball(const MaskBuffers &B, const int64_t *pi, const int64_t *pj, int64_t npairs, double *lb,
                       hipStream_t s) {
    if (npairs <= 0) return 0;
    MaskArgs A{B.a, B.U, B.L, B.mu, B.sig, B.n, B.n - B.m + 1, B.m, B.w, (int32_t)B.lmax};
    hipLaunchKernelGGL(k_masked_ball, dim3((unsigned)npairs), dim3(64), 0, s, A, (const uint8_t *)B.mask, pi, pj,
                       npairs, lb);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

}  // namespace dtwj
