// ED matrix profile on gfx950 (reference: mpx_v2, GPU_version1/src/mpx_v2.cu:14-150 and
// CPU_version1/mpx_v2.cpp:13-150, called with minlag = m by DTWMotifDiscoveryGUI): for every
// window t, the z-normalised Euclidean distance to its nearest window with |t - p| >= m and the
// 1-based index of that window.  It seeds the reference's join (the 100 smallest entries,
// GPU_version1/src/new_dtw_motifGUI_malloc.cu:82-107).
//
// All arithmetic is FP64.  Per window: mean and population standard deviation by two passes,
// and the MPX update terms df[t] = (a[t+m] - a[t]) / 2, dg[t] = (a[t+m] - mu[t+1]) + (a[t] - mu[t]),
// with which the centred covariance of a diagonal moves by one row in two FMAs:
//     cov(i+1, j+1) = cov(i, j) + df[i] dg[j] + df[j] dg[i].
// ED^2(i, j) = 2m - 2 cov(i, j) / (sigma_i sigma_j)  (= 2m (1 - rho), the reference's
// sqrt(2 m (1 - corr)) with corr = cov / (||a_i - mu_i|| ||a_j - mu_j||)).
//
// Sweep: a task is (row block of R rows, G consecutive groups of 256 diagonal offsets); one
// wave sweeps the task's groups one after the other.  Per group the covariances of the lane's
// four diagonals are anchored exactly at the first row (windows staged in LDS), then carried
// down the rows with the update above; the column terms rotate through registers (slot t of a
// lane holds column r0 + d0 + 4 lane + t of a 4-row block starting at row r0; the four columns
// entering a block are read by each lane from memory, where they are L1 hits).  Each row's minimum over
// the task's diagonals goes into an LDS array of row keys (FP32 bits of ED^2, the lowest two
// cleared, << 32 | partner: smaller distance first, then the smaller partner index), flushed to
// the profile with one 64-bit atomicMin per row and task.  One sweep gives every window its nearest partner to the
// right (p >= t + m); the same sweep over the reversed series gives the partners to the left.
// A final pass merges both and recomputes the distance of the chosen pair exactly by a direct
// FP64 dot product.
//
// Tie rule: partners whose ED^2 agree in the FP32 bits kept (22 of the 23 fraction bits: a
// relative difference below 2^-21) are tied, and the one with the
// smallest lag |t - p| is taken, the left one (p < t) at equal lags: the partner the reference
// meets first in its diagonal order among exactly equal correlations (CPU_version1/mpx_v2.cpp:
// diagonals in increasing lag, and at one lag window t as the column before t as the row).  The
// reported distance is the chosen pair's own FP64 distance.
#include "../../include/dtwj.h"
#include "internal.hpp"

namespace dtwj {

namespace {

constexpr int MP_KD = 4;               // diagonals per lane
constexpr int MP_DG = 64 * MP_KD;      // diagonals per group
constexpr int MP_ROWS = 1024;          // rows per task
constexpr int MP_GPT = 32;             // groups per task

// Window statistics, one thread per window: mean, 1/sigma (two passes), and a flag for a
// constant window (sigma = 0: the profile is undefined there, as the join's input is).
__global__ __launch_bounds__(256) void k_mp_stats(const double *__restrict__ a, int64_t nw, int m, double *mu,
                                                  double *isig, int32_t *flags) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nw) return;
    const double *x = a + t;
    double s = 0.0;
    for (int k = 0; k < m; ++k) s += x[k];
    const double mean = s / m;
    double q = 0.0;
    for (int k = 0; k < m; ++k) {
        const double d = x[k] - mean;
        q = fma(d, d, q);
    }
    const double sig = sqrt(q / m);
    mu[t] = mean;
    isig[t] = sig > 0.0 ? 1.0 / sig : 0.0;
    if (!(sig > 0.0)) atomicOr(&flags[1], 1);
    if (!isfinite(s) || !isfinite(q)) atomicOr(&flags[0], 1);
}

// MPX update terms of moving window t to t + 1 (0 for the last window).
__global__ __launch_bounds__(256) void k_mp_terms(const double *__restrict__ a, int64_t nw, int m,
                                                  const double *__restrict__ mu, double *df, double *dg) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nw) return;
    if (t + 1 < nw) {
        const double am = a[t + m], a0 = a[t];
        df[t] = 0.5 * (am - a0);
        dg[t] = (am - mu[t + 1]) + (a0 - mu[t]);
    } else {
        df[t] = 0.0;
        dg[t] = 0.0;
    }
}

__global__ __launch_bounds__(256) void k_mp_reverse(const double *__restrict__ a, int64_t n, double *r) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) r[t] = a[n - 1 - t];
}

__global__ __launch_bounds__(256) void k_mp_fill(unsigned long long *key, int64_t nw, int32_t *flags) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < nw) key[t] = ~0ull;
    if (t < 2 && flags) flags[t] = 0;
}

struct MpSweep {
    const double *a;
    const double *mu, *isig, *df, *dg;
    unsigned long long *key;
    int64_t n, nw;
    int32_t m, groups;
};

// One task per block: rows [blockIdx.x R, + R), groups [blockIdx.y G, + G).  Pairs (i, j = i + d)
// with d >= m and j <= nw - 1.  LDS: R row keys, then the anchor's staged windows.
__global__ __launch_bounds__(64) void k_mpx(MpSweep S) {
    extern __shared__ __align__(16) double mp_lds[];
    const int lane = (int)threadIdx.x;
    const int64_t nw = S.nw, n = S.n;
    const int m = S.m;
    const int64_t i0 = (int64_t)blockIdx.x * MP_ROWS;
    const int g0 = (int)blockIdx.y * MP_GPT;
    if (i0 + m + (int64_t)MP_DG * g0 > nw - 1) return;  // the task holds no pair
    unsigned long long *rk = reinterpret_cast<unsigned long long *>(mp_lds);
    double *lrow = mp_lds + MP_ROWS;  // m values: the row window, centred
    double *lcol = lrow + m;          // MP_DG + m + MP_KD values from column i0 + d0
    for (int x = lane; x < MP_ROWS; x += 64) rk[x] = ~0ull;
    const int gend = g0 + MP_GPT < S.groups ? g0 + MP_GPT : S.groups;
    for (int g = g0; g < gend; ++g) {
        const int64_t d0 = (int64_t)m + (int64_t)MP_DG * g;
        if (i0 + d0 > nw - 1) break;  // this group (and every later one) has no row here
        const int64_t i_end = (i0 + MP_ROWS < nw - d0) ? i0 + MP_ROWS : nw - d0;
        const int64_t jl = i0 + d0 + MP_KD * lane;  // the lane's first column at row i0
        // ---- exact anchor at row i0: cov_q = sum_k (a[i0+k] - mu[i0]) (a[j_q+k] - mu[j_q])
        __syncthreads();  // the previous group's staging is no longer read
        {
            const double mi = S.mu[i0];
            for (int x = lane; x < m; x += 64) lrow[x] = S.a[i0 + x] - mi;
            const int span = MP_DG + m + MP_KD;
            for (int x = lane; x < span; x += 64) {
                const int64_t p = i0 + d0 + x;
                lcol[x] = p < n ? S.a[p] : 0.0;
            }
        }
        __syncthreads();
        double cov[MP_KD];
        {
            double acc[MP_KD] = {0.0, 0.0, 0.0, 0.0}, rsum = 0.0;
            const double *cl = lcol + MP_KD * lane;
            double win[MP_KD];
#pragma unroll
            for (int q = 0; q < MP_KD - 1; ++q) win[q + 1] = cl[q];
#pragma unroll 4
            for (int k = 0; k < m; ++k) {
#pragma unroll
                for (int q = 0; q < MP_KD - 1; ++q) win[q] = win[q + 1];
                win[MP_KD - 1] = cl[k + MP_KD - 1];
                const double ck = lrow[k];
                rsum += ck;
#pragma unroll
                for (int q = 0; q < MP_KD; ++q) acc[q] = fma(ck, win[q], acc[q]);
            }
            // sum_k (a_i - mu_i) a_j = cov + mu_j sum_k (a_i - mu_i): remove the rounding residue
#pragma unroll
            for (int q = 0; q < MP_KD; ++q) {
                const int64_t j = jl + q;
                cov[q] = j <= nw - 1 ? acc[q] - S.mu[j] * rsum : 0.0;
            }
        }
        // ---- sweep: 4-row blocks; slots t = 0..7 hold columns r0 + d0 + 4 lane + t
        double cdf[8], cdg[8], cis[8];
        auto load4 = [&](int base_t, int64_t r0) __attribute__((always_inline)) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                int64_t j = r0 + d0 + MP_KD * lane + base_t + t;
                j = j < nw - 1 ? j : nw - 1;  // columns past the last window only feed masked pairs
                cdf[base_t + t] = S.df[j];
                cdg[base_t + t] = S.dg[j];
                cis[base_t + t] = S.isig[j];
            }
        };
        load4(0, i0);
        const double m2 = 2.0 * (double)m;
        for (int64_t r0 = i0; r0 < i_end; r0 += 4) {
            load4(4, r0);
            double rdf[4], rdg[4], ris[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                int64_t i = r0 + u;
                i = i < nw - 1 ? i : nw - 1;
                rdf[u] = S.df[i];
                rdg[u] = S.dg[i];
                ris[u] = S.isig[i];
            }
            // The block's four rows.  Each distance becomes a key: its FP32 bits with the two lowest
            // bits replaced by the diagonal q of the lane, so that one FP32 minimum per pair keeps the
            // smallest distance and, among distances equal in the other 30 bits, the smallest q
            // (the partner index is rebuilt from the winning lane and q).  Pairs past the last
            // window (only in blocks at the triangle's edge) get +inf.
            const bool full = r0 + 3 < i_end && r0 + 3 + d0 + MP_DG - 1 <= nw - 1;  // uniform
            const int64_t jl0 = d0 + MP_KD * lane;  // j - i of the lane's first diagonal
            float best[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t i = r0 + u;
                const double s2 = -2.0 * ris[u];
                best[u] = __builtin_huge_valf();
#pragma unroll
                for (int q = 0; q < MP_KD; ++q) {
                    const double e2 = fma(s2, cov[q] * cis[u + q], m2);
                    const float e = fmaxf((float)e2, 0.f);
                    float key = __uint_as_float((__float_as_uint(e) & ~3u) | (uint32_t)q);
                    if (!full && !(i < i_end && i + jl0 + q <= nw - 1)) key = __builtin_huge_valf();
                    best[u] = fminf(best[u], key);
                    cov[q] = fma(rdf[u], cdg[u + q], cov[q]);
                    cov[q] = fma(cdf[u + q], rdg[u], cov[q]);
                }
            }
            float v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = best[u];
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = fminf(v[u], __shfl_xor(v[u], off, 64));
            }
            // the lowest lane holding the minimum key has the smallest partner among equal keys
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (!(v[u] < __builtin_huge_valf())) continue;  // uniform: no valid pair in the row
                const unsigned long long hit = __builtin_amdgcn_ballot_w64(best[u] == v[u]);
                const int src = (int)__builtin_ctzll(hit);
                const uint32_t kb = __float_as_uint(v[u]);
                const int64_t jj = r0 + u + d0 + MP_KD * src + (kb & 3u);
                if (lane == 0)
                    atomicMin(&rk[r0 + u - i0], ((unsigned long long)(kb & ~3u) << 32) | (unsigned long long)(uint32_t)jj);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                cdf[t] = cdf[t + 4];
                cdg[t] = cdg[t + 4];
                cis[t] = cis[t + 4];
            }
        }
    }
    __syncthreads();
    for (int x = lane; x < MP_ROWS; x += 64) {
        const unsigned long long k = rk[x];
        if (k != ~0ull && i0 + x < nw) atomicMin(&S.key[i0 + x], k);
    }
}

// Merges the right-hand partners (key_f, series order) and the left-hand ones (key_r, found on
// the reversed series: its window t' is window nw - 1 - t' here) and writes the profile: the
// chosen pair's FP64 distance and the partner's 1-based index.  A window with no partner at
// least m away gets sqrt(4m) and index 0 (the reference's initial correlation -1).
__global__ __launch_bounds__(256) void k_mp_finish(const double *__restrict__ a, int64_t nw, int m,
                                                   const double *__restrict__ mu, const double *__restrict__ isig,
                                                   const unsigned long long *__restrict__ key_f,
                                                   const unsigned long long *__restrict__ key_r, double *mp,
                                                   int32_t *mpi) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nw) return;
    unsigned long long kf = key_f[t];
    unsigned long long kr = key_r[nw - 1 - t];
    if (kr != ~0ull) {  // map the reversed partner index back to this series
        const uint32_t p = (uint32_t)(nw - 1 - (int64_t)(uint32_t)kr);
        kr = (kr & 0xffffffff00000000ull) | p;
    }
    // equal FP32 distances: the smaller lag, the left partner at equal lags (each sweep already
    // kept its smallest lag: the right sweep the smallest index, the reversed one the largest)
    unsigned long long k = kf;
    if (kr != ~0ull) {
        const uint32_t vf = (uint32_t)(kf >> 32), vr = (uint32_t)(kr >> 32);
        const int64_t lf = kf == ~0ull ? INT64_MAX : (int64_t)(uint32_t)kf - t;
        const int64_t lr = t - (int64_t)(uint32_t)kr;
        if (kf == ~0ull || vr < vf || (vr == vf && lr <= lf)) k = kr;
    }
    if (k == ~0ull) {
        mp[t] = sqrt(4.0 * m);
        mpi[t] = 0;
        return;
    }
    const int64_t p = (int64_t)(uint32_t)k;
    const double *x = a + t, *y = a + p;
    const double mx = mu[t], my = mu[p];
    double c = 0.0;
    for (int i = 0; i < m; ++i) c = fma(x[i] - mx, y[i] - my, c);
    const double e2 = 2.0 * m - 2.0 * c * isig[t] * isig[p];
    mp[t] = sqrt(e2 > 0.0 ? e2 : 0.0);
    mpi[t] = (int32_t)(p + 1);
}

}  // namespace

size_t mp_lds_bytes(int32_t m) { return sizeof(double) * ((size_t)MP_ROWS + 2 * (size_t)m + MP_DG + MP_KD); }

int64_t mp_scratch_doubles(int64_t n, int32_t m) {
    const int64_t nw = n - m + 1;
    // reversed series (n), per direction: mu, isig, df, dg (nw each) and keys (nw)
    return n + 2 * 5 * nw;
}

int launch_matrix_profile(const double *d_a, int64_t n, int32_t m, double *scratch, int32_t *flags, double *d_mp,
                          int32_t *d_mpi, hipStream_t s) {
    const int64_t nw = n - m + 1;
    const int64_t K = nw - m;  // diagonal offsets m .. nw - 1
    if (K <= 0) {
        set_error("series too short for a matrix profile: need n >= 2m");
        return DTWJ_EINVAL;
    }
    const size_t lds = mp_lds_bytes(m);
    if (lds > 64 * 1024) {
        set_error("window too long for the matrix profile's LDS staging");
        return DTWJ_EINVAL;
    }
    double *ar = scratch;
    double *base = ar + n;
    double *mu[2], *is[2], *df[2], *dg[2];
    unsigned long long *key[2];
    for (int d = 0; d < 2; ++d) {
        mu[d] = base + (size_t)(5 * d + 0) * nw;
        is[d] = base + (size_t)(5 * d + 1) * nw;
        df[d] = base + (size_t)(5 * d + 2) * nw;
        dg[d] = base + (size_t)(5 * d + 3) * nw;
        key[d] = reinterpret_cast<unsigned long long *>(base + (size_t)(5 * d + 4) * nw);
    }
    const unsigned bn = (unsigned)((n + 255) / 256), bw = (unsigned)((nw + 255) / 256);
    hipLaunchKernelGGL(k_mp_reverse, dim3(bn), dim3(256), 0, s, d_a, n, ar);
    const int32_t groups = (int32_t)((K + MP_DG - 1) / MP_DG);
    const dim3 grid((unsigned)((nw + MP_ROWS - 1) / MP_ROWS), (unsigned)((groups + MP_GPT - 1) / MP_GPT));
    // flags (non-finite sample, constant window) are cleared by the first key fill
    hipLaunchKernelGGL(k_mp_fill, dim3(bw), dim3(256), 0, s, key[0], nw, flags);
    hipLaunchKernelGGL(k_mp_fill, dim3(bw), dim3(256), 0, s, key[1], nw, (int32_t *)nullptr);
    for (int d = 0; d < 2; ++d) {
        const double *src = d == 0 ? d_a : ar;
        hipLaunchKernelGGL(k_mp_stats, dim3(bw), dim3(256), 0, s, src, nw, m, mu[d], is[d], flags);
        hipLaunchKernelGGL(k_mp_terms, dim3(bw), dim3(256), 0, s, src, nw, m, mu[d], df[d], dg[d]);
        MpSweep S{src, mu[d], is[d], df[d], dg[d], key[d], n, nw, m, groups};
        hipLaunchKernelGGL(k_mpx, grid, dim3(64), lds, s, S);
    }
    hipLaunchKernelGGL(k_mp_finish, dim3(bw), dim3(256), 0, s, d_a, nw, m, mu[0], is[0], key[0], key[1], d_mp,
                       d_mpi);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

}  // namespace dtwj
