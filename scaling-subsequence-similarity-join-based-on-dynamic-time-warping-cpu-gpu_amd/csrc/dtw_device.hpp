// Banded DTW of one pair on one 64-lane wavefront (device code).
//
// Cells are visited by anti-diagonals t = r + c.  Lane l owns the band offsets
// delta = c - r = l + 64 q - w for q = 0..Q-1 (Q = ceil((2w+1)/64) is a template parameter,
// so any band up to 64*Q-1 cells runs without recompiling; it replaces the reference's
// compile-time REGISTER_NUM, GPU_version1/src/DTW_stairs.cu:5).  A lane keeps one value per
// owned offset: the cell of the latest anti-diagonal of that offset's parity.  Cell (r, c)
// on anti-diagonal t needs (r-1, c-1) [same offset, t-2: own register], (r-1, c) [offset+1,
// t-1: right neighbour] and (r, c-1) [offset-1, t-1: left neighbour]; neighbours arrive by
// one-lane wave shifts.  Recurrence as in the reference kernels
// (GPU_version1/src/DTW_stairs.cu:446-599): D = (x_r - y_c)^2 + min(diag, up, left),
// out-of-band cells infinite.
//
// Early abandon: every path crosses anti-diagonal t or t-1, so once the minimum over all
// lanes' registers exceeds ub the result can only exceed ub; the routine then returns +inf
// (the reference abandons with a ballot at the band's middle, DTW_stairs.cu:225-237).
#pragma once

#include "common.hpp"

namespace dtwj {

template <typename T>
__device__ inline T dev_inf();
template <>
__device__ inline float dev_inf<float>() {
    return __builtin_huge_valf();
}
template <>
__device__ inline double dev_inf<double>() {
    return __builtin_huge_val();
}

template <typename T>
__device__ inline T wave_min(T x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T y = __shfl_xor(x, off, 64);
        x = y < x ? y : x;
    }
    return x;
}

// zi, zj: m values each (LDS).  Returns the squared cost D[m-1][m-1], or +inf when abandoned.
// cells: number of cells evaluated (uniform).  Must be called by all 64 lanes.
template <typename T, int Q>
__device__ T dtw_wave(const T *zi, const T *zj, int m, int w, T ub, int64_t &cells, bool &abandoned) {
    const int lane = (int)(threadIdx.x & 63);
    const T INF = dev_inf<T>();
    T D[Q];
    int dl[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        dl[q] = lane + 64 * q - w;
        D[q] = INF;
        if (dl[q] == 0) {
            T d = zi[0] - zj[0];
            D[q] = d * d;
        }
    }
    cells = 1;
    abandoned = false;
    const int t_end = 2 * m - 2;
    for (int t = 1; t <= t_end; ++t) {
        T up[Q], lf[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            T dn = __shfl_down(D[q], 1, 64);  // lane l receives lane l+1
            T uu = __shfl_up(D[q], 1, 64);    // lane l receives lane l-1
            if (Q > 1) {
                T from0 = __shfl(D[(q + 1 < Q) ? q + 1 : q], 0, 64);
                T from63 = __shfl(D[(q > 0) ? q - 1 : q], 63, 64);
                up[q] = (lane == 63) ? ((q + 1 < Q) ? from0 : INF) : dn;
                lf[q] = (lane == 0) ? ((q > 0) ? from63 : INF) : uu;
            } else {
                up[q] = (lane == 63) ? INF : dn;
                lf[q] = (lane == 0) ? INF : uu;
            }
        }
        int active = 0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int dd = dl[q];
            if (((t - dd) & 1) == 0 && dd >= -w && dd <= w) {
                const int r = (t - dd) >> 1;
                const int c = r + dd;
                if (r >= 0 && c >= 0 && r < m && c < m) {
                    T d = zi[r] - zj[c];
                    T best = D[q] < up[q] ? D[q] : up[q];
                    best = best < lf[q] ? best : lf[q];
                    D[q] = d * d + best;
                    ++active;
                } else {
                    D[q] = INF;
                }
            }
        }
        // count evaluated cells (uniform)
        unsigned long long b1 = __ballot(active >= 1);
        cells += (int64_t)__popcll(b1);
        if (Q > 1) {
            for (int k = 2; k <= Q; ++k) cells += (int64_t)__popcll(__ballot(active >= k));
        }
        if ((t & 15) == 0) {
            T mn = D[0];
#pragma unroll
            for (int q = 1; q < Q; ++q) mn = D[q] < mn ? D[q] : mn;
            mn = wave_min(mn);
            if (mn > ub) {
                abandoned = true;
                return INF;
            }
        }
    }
    // result: offset 0 lives in lane w % 64, register w / 64
    const int qsel = w >> 6;
    T res = INF;
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (q == qsel) res = D[q];
    res = __shfl(res, w & 63, 64);
    if (res > ub) {
        abandoned = true;
        return INF;
    }
    return res;
}

// Lanes per band: Q such that 2w+1 <= 64*Q.
inline int dtw_q_for(int w) {
    int band = 2 * w + 1;
    if (band <= 64) return 1;
    if (band <= 128) return 2;
    if (band <= 256) return 4;
    if (band <= 512) return 8;
    if (band <= 1024) return 16;
    return 0;
}

}  // namespace dtwj
