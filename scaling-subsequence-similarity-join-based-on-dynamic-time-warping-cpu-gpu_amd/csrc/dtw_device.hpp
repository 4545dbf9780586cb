// Banded DTW on one 64-lane wavefront (device code).
//
// Cells are visited by anti-diagonals t = r + c, t = 0 .. 2m-2.  A cell's band offset is
// delta = c - r, |delta| <= w, and the cells of anti-diagonal t have delta = t (mod 2).  Each
// lane owns QP pairs of adjacent offsets (e, o = e + 1), e = 2 (x QP + k) - w for lane x of its
// segment and k = 0..QP-1; every step therefore evaluates exactly one cell per pair on every
// lane: the even-parity offsets on steps with t = w (mod 2), the odd ones on the others.
// Recurrence (as the reference kernels, GPU_version1/src/DTW_stairs.cu:446-599):
//   D(r, c) = (x_r - y_c)^2 + min(D(r-1, c-1), D(r-1, c), D(r, c-1)),   D(-1, -1) = 0,
// with (r-1, c-1) = same offset two steps back (own register), (r-1, c) = offset + 1 one step
// back and (r, c-1) = offset - 1 one step back: own registers except across the lane edge,
// which arrives by one DPP wave shift per step.  QP is a template parameter, so any band up to
// 128 QP offsets runs without recompiling (it replaces the compile-time REGISTER_NUM of
// GPU_version1/src/DTW_stairs.cu:5).  Offsets beyond w are forced to +inf.
//
// With QP = 1 and w <= 63 the wave holds several pairs: segments of seg_len >= w + 1 lanes, one
// pair per segment; the last lane of a segment owns an out-of-band offset, so the shifts never
// carry values between pairs.
//
// Early abandon: every path crosses anti-diagonal t or t+1 (the two held in the registers at a
// check), so once the minimum over a segment's registers exceeds ub its result can only exceed
// ub; the segment is then reported abandoned (the reference abandons with a ballot at the
// band's middle, DTW_stairs.cu:225-237).  Optionally a suffix bound is added to the minimum:
// after its crossing cell (r, c), r + c <= t + 1 and |r - c| <= w give r, c <= (t + 1 + w) / 2,
// and the rest of the path visits every later row and every later column, each at a cost of at
// least its LB_Keogh term (the cumulative-bound early abandon of the UCR suite, applied per
// anti-diagonal check).  The suffix functor returns that sum for the check at t.
//
// FP32 safety of the suffix test: the envelope values are samples (stored with the same
// rounding), and the caller forms each term with the cells' own arithmetic, which is monotone
// in every sample value, so each term is at most the FP32 cell cost it stands for; the FMA
// chain of the remaining path loses at most (2m + 2) u of
// relative value and the FP32 suffix sum is at most (m + 1) u too large (u = 2^-24).  Scaling
// (minimum + suffix) by (1 - 2^-9) therefore never abandons a pair whose FP32 result would be
// <= ub, for m up to 8192 (3 (m + 2) u <= 2^-9).
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

// Lane L receives src of lane L - 1 (lane 0: edge) / of lane L + 1 (lane 63: edge): DPP wave
// shifts, the out-of-range lane keeping `edge`.
__device__ inline uint32_t dpp_from_left_u32(uint32_t edge, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)src, 0x138, 0xf, 0xf, false);
}
__device__ inline uint32_t dpp_from_right_u32(uint32_t edge, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)src, 0x130, 0xf, 0xf, false);
}
template <typename T>
__device__ inline T from_left(T edge, T src);
template <typename T>
__device__ inline T from_right(T edge, T src);
template <>
__device__ inline float from_left<float>(float edge, float src) {
    return __builtin_bit_cast(float, dpp_from_left_u32(__builtin_bit_cast(uint32_t, edge), __builtin_bit_cast(uint32_t, src)));
}
template <>
__device__ inline float from_right<float>(float edge, float src) {
    return __builtin_bit_cast(float, dpp_from_right_u32(__builtin_bit_cast(uint32_t, edge), __builtin_bit_cast(uint32_t, src)));
}
template <>
__device__ inline double from_left<double>(double edge, double src) {
    const uint64_t e = __builtin_bit_cast(uint64_t, edge), s = __builtin_bit_cast(uint64_t, src);
    const uint32_t lo = dpp_from_left_u32((uint32_t)e, (uint32_t)s);
    const uint32_t hi = dpp_from_left_u32((uint32_t)(e >> 32), (uint32_t)(s >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <>
__device__ inline double from_right<double>(double edge, double src) {
    const uint64_t e = __builtin_bit_cast(uint64_t, edge), s = __builtin_bit_cast(uint64_t, src);
    const uint32_t lo = dpp_from_right_u32((uint32_t)e, (uint32_t)s);
    const uint32_t hi = dpp_from_right_u32((uint32_t)(e >> 32), (uint32_t)(s >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ inline float min3(float a, float b, float c) { return fminf(fminf(a, b), c); }
__device__ inline double min3(double a, double b, double c) { return fmin(fmin(a, b), c); }

// z-values of a DTW pair already normalised in memory (indices must lie in [0, m-1]).
template <typename T>
struct PlainZ {
    static constexpr bool kPadded = false;
    static constexpr bool kFused = false;
    const T *zi, *zj;
    __device__ T row(int r) const { return zi[r]; }
    __device__ T col(int c) const { return zj[c]; }
};

// No suffix bound: the abandon test uses the partial minimum alone.
struct NoSuffix {
    static constexpr bool enabled = false;
    __device__ float at(int) const { return 0.f; }
};

// Cells (r, c) with r + c <= t, |r - c| <= w, 0 <= r, c <= m - 1 (w <= m - 1): the cells a
// DTW evaluated when it stopped after anti-diagonal t.
__host__ __device__ inline int64_t dtw_cells_upto(int m, int w, int64_t t) {
    auto G = [w](int64_t s) -> int64_t {  // same count without the upper bounds
        if (s < 0) return 0;
        if (s <= w) return (s + 1) * (s + 2) / 2;
        return (int64_t)(w + 1) * (w + 2) / 2 + (s - w) * w + (s - w) / 2;
    };
    const int64_t total = (int64_t)m * (2 * w + 1) - (int64_t)w * (w + 1);
    if (t >= 2 * (int64_t)m - 2) return total;
    if (t <= m - 1) return G(t);
    return total - G(2 * (int64_t)m - 3 - t);
}

// Offset pairs per lane for band radius w (one pair per wave): 128 QP >= 2w + 1.
inline int dtw_qp_for(int w) {
    const int band = 2 * w + 1;
    if (band <= 128) return 1;
    if (band <= 256) return 2;
    if (band <= 512) return 4;
    if (band <= 1024) return 8;
    return 0;
}

// dtw_band for a Z of raw staged values (Z::kFused, e.g. the join's StagedZ): a cell's
// difference is formed as d = fma(x_c, z.nij, fma(x_r, z.ii, c)), i.e. z_i(r) - z_j(c) with
// c = mu_j / sigma_j - mu_i / sigma_i, two FMAs per pair of cells sharing a row and no
// separate normalisation of the columns.  The row term of an offset pair wholly past the band
// (both members out, or a lane outside every segment) takes c = +inf: its differences are then
// +inf and its cells +inf without a select; the odd member of the pair whose even member is
// the band's last offset w takes a NaN column coefficient instead (nijo).  Same contract and
// results as dtw_band otherwise.
template <typename T, int QP, typename Z, typename SUF>
__device__ __forceinline__ T dtw_band_fused(int m, int w, int seg_len, int nseg, T ub, const Z &z, int &t_stop,
                                            SUF suf) {
    const int lane = (int)(threadIdx.x & 63);
    const T INF = dev_inf<T>();
    const int seg = lane / seg_len;
    const int x = lane - seg * seg_len;
    const bool valid = seg < nseg;
    // lanes of this lane's segment
    const unsigned long long segmask =
        valid ? (seg_len >= 64 ? ~0ull : ((1ull << seg_len) - 1ull) << (seg * seg_len)) : 0ull;
    T E[QP], O[QP];
    int e_off[QP];
    bool inE[QP], inO[QP];
    // the slot of offset 0 holds D(-1, -1) = 0 before the first step
    const int idx0 = w >> 1;              // pair index of offset 0
    const bool zero_even = (w & 1) == 0;  // offset 0 is the even member of its pair
#pragma unroll
    for (int k = 0; k < QP; ++k) {
        const int xk = x * QP + k;
        e_off[k] = 2 * xk - w;
        inE[k] = valid && e_off[k] <= w;
        inO[k] = valid && e_off[k] + 1 <= w;
        E[k] = (valid && xk == idx0 && zero_even) ? (T)0 : INF;
        O[k] = (valid && xk == idx0 && !zero_even) ? (T)0 : INF;
    }
    const int mm1 = m - 1;
    const int t_end = 2 * m - 2;
    auto clampi = [mm1](int v) {
        if constexpr (Z::kPadded)
            return v;
        else
            return min(max(v, 0), mm1);
    };
    // offsets used for the z-value indices: lanes past the band (their cells are forced to
    // +inf) read at offset w + 1, so that the indices stay within the padded range
    int e_ix[QP];
#pragma unroll
    for (int k = 0; k < QP; ++k) e_ix[k] = min(e_off[k], w + 1);
    static_assert(Z::kFused, "dtw_band_fused takes raw staged values");
    // row bias of each offset pair: +inf for pairs wholly past the band
    T cbias[QP];
#pragma unroll
    for (int k = 0; k < QP; ++k) cbias[k] = inE[k] ? (T)z.c : INF;
    auto rowv = [&](int rr) __attribute__((always_inline)) { return (T)z.xrow(clampi(rr)); };
    auto colv = [&](int cc) __attribute__((always_inline)) { return (T)z.xcol(clampi(cc)); };
    // row term of pair k (shared by its even and odd cell) and a cell's difference
    auto rowz = [&](int k, T xr) __attribute__((always_inline)) { return (T)fma(xr, (T)z.ii, cbias[k]); };
    auto diff = [&](T zr, T xc) __attribute__((always_inline)) { return (T)fma(xc, (T)z.nij, zr); };
    // column coefficient of each odd cell: z.nij, or NaN for an odd member past the band.  Its
    // difference and cell are then NaN, which every min3 here ignores as it would +inf (IEEE
    // minNum), and which no check sees (a check takes min3 from +inf); so no select per cell.
    T nijo[QP];
#pragma unroll
    for (int k = 0; k < QP; ++k) nijo[k] = inO[k] ? (T)z.nij : (T)__builtin_nan("");
    auto diffo = [&](T zr, T xc, int k) __attribute__((always_inline)) { return (T)fma(xc, nijo[k], zr); };
    // neighbour values across lanes (DPP wave shifts): the edge lane of each shift (lane 0 from
    // the left, lane 63 from the right) has no source and keeps the register's old value, which
    // is +inf from here on, so the shifts need no +inf reloaded each step
    T lin = INF, rin = INF;
    auto odd_step = [&](int r[QP], T zr[QP]) __attribute__((always_inline)) {
        // odd offsets: up = next even (t-1), left = own even (t-1), diag = own odd (t-2)
        rin = from_right<T>(rin, E[0]);
        const T right_in = rin;
#pragma unroll
        for (int k = 0; k < QP; ++k) {
            const T up = k + 1 < QP ? E[k + 1] : right_in;
            const T d = diffo(rowz(k, zr[k]), colv(r[k] + e_ix[k] + 1), k);
            O[k] = fma(d, d, min3(O[k], up, E[k]));
        }
    };
    bool alive = valid;
    t_stop = t_end;
    int t = 0;
    int r[QP];
    T zr[QP];  // raw row values
    if (!zero_even) {
        // t = 0 has odd parity: only the odd offsets, on row (0 - e - 1) / 2
#pragma unroll
        for (int k = 0; k < QP; ++k) {
            r[k] = (0 - e_ix[k] - 1) >> 1;
            zr[k] = rowv(r[k]);
        }
        odd_step(r, zr);
        t = 1;
    }
    // values are read one iteration ahead (their latency overlaps a whole iteration): the
    // row value of the pair of steps, and the two columns r + e and r + e + 1; the odd step's
    // column is the next iteration's even-step column.
    T zce[QP], zco[QP];
#pragma unroll
    for (int k = 0; k < QP; ++k) {
        r[k] = (t - e_ix[k]) >> 1;
        zr[k] = rowv(r[k]);
        zce[k] = colv(r[k] + e_ix[k]);
        zco[k] = colv(r[k] + e_ix[k] + 1);
    }
    // One pair of steps: the even offsets at t and, when `odd`, the odd offsets at t + 1.
    auto step_pair = [&](bool odd) __attribute__((always_inline)) {
        T zr_n[QP], zco_n[QP];
#pragma unroll
        for (int k = 0; k < QP; ++k) {
            zr_n[k] = rowv(r[k] + 1);
            zco_n[k] = colv(r[k] + e_ix[k] + 2);
        }
        T rz[QP];
#pragma unroll
        for (int k = 0; k < QP; ++k) rz[k] = rowz(k, zr[k]);
        // even offsets at t: up = own odd (t-1), left = previous odd (t-1), diag = own even (t-2)
        lin = from_left<T>(lin, O[QP - 1]);
        const T left_in = lin;
#pragma unroll
        for (int k = 0; k < QP; ++k) {
            const T left = k > 0 ? O[k - 1] : left_in;
            const T d = diff(rz[k], zce[k]);
            E[k] = fma(d, d, min3(E[k], O[k], left));  // +inf past the band through the row term
        }
        if (odd) {
            // odd offsets: up = next even (t), left = own even (t), diag = own odd (t-1)
            rin = from_right<T>(rin, E[0]);
            const T right_in = rin;
#pragma unroll
            for (int k = 0; k < QP; ++k) {
                const T up = k + 1 < QP ? E[k + 1] : right_in;
                const T d = diffo(rz[k], zco[k], k);
                O[k] = fma(d, d, min3(O[k], up, E[k]));
            }
        }
#pragma unroll
        for (int k = 0; k < QP; ++k) {
            ++r[k];
            zr[k] = zr_n[k];
            zce[k] = zco[k];
            zco[k] = zco_n[k];
        }
    };
    // Abandon test after the pair of steps at tc (tc = 14 or 15 mod 16); true when no segment
    // of the wave is left.  The segment is over ub when its minimum over both live
    // anti-diagonals is; both tests are monotone in the value, so that holds exactly when every
    // lane of the segment is over on its own cells: one ballot instead of a cross-lane minimum.
    auto check = [&](int tc) __attribute__((always_inline)) {
        T v = INF;
#pragma unroll
        for (int k = 0; k < QP; ++k) v = min3(v, E[k], O[k]);
        bool over = v > ub;
        if constexpr (SUF::enabled) {
            const T sb = (T)suf.at(tc);  // every lane takes part (the functor may shuffle)
            over = over || (v + sb) * (T)(1.0 - 0x1p-9) > ub;
        }
        const unsigned long long keep = __builtin_amdgcn_ballot_w64(!over);
        if (alive && (keep & segmask) == 0ull) {
            alive = false;
            t_stop = tc + 1;
        }
        return __builtin_amdgcn_ballot_w64(alive) == 0ull;
    };
    // Blocks of 8 pairs of steps (t = t0 + 16 k .. + 15, t0 = t mod 2 here) end in a test, as
    // long as the block's last odd step t + 15 comes before the last anti-diagonal; in the
    // remaining tail (t + 15 >= t_end) no test falls, since a test needs tc + 1 < t_end.
    bool done = false;
    while (t + 15 < t_end) {
#pragma unroll
        for (int u = 0; u < 8; ++u) step_pair(true);
        t += 16;
        if (check(t - 2)) {
            done = true;
            break;
        }
    }
    if (!done)
        for (; t <= t_end; t += 2) step_pair(t + 1 <= t_end);
    // result: offset 0 of the lane's segment
    const int src_lane = valid ? seg * seg_len + idx0 / QP : lane;
    const int k0 = idx0 % QP;
    T mine = INF;
#pragma unroll
    for (int k = 0; k < QP; ++k)
        if (k == k0) mine = zero_even ? E[k] : O[k];
    const T res = __shfl(mine, src_lane, 64);
    if (!alive || !valid) return INF;
    if (res > ub) {
        t_stop = t_end;
        return INF;
    }
    return res;
}

// Banded DTW of up to `nseg` pairs at once (QP = 1: segments of seg_len lanes, seg_len >= w+1;
// QP > 1: nseg = 1, seg_len = 64).  z.row(r) / z.col(c) give the z-values of the lane's pair.
// Without Z::kPadded they are only called with indices clamped into [0, m-1]; with it, with
// indices in [-(w+3)/2 - 1, m + (w+3)/2 + 1], where they must return finite values (cells
// outside the matrix then take arbitrary finite costs: those before it stay +inf, as all their
// predecessors are, and those after it feed no cell inside it).  Returns, on every lane, the
// squared cost D(m-1, m-1) of the lane's pair, or +inf when the pair was abandoned against ub
// (then t_stop is the last anti-diagonal evaluated, else 2m-2).
//
// Steps come in pairs (even offsets at t, odd offsets at t+1): both cells of a lane's offset
// pair lie on the same row r = (t - e)/2, so a pair of steps reads one row value and two
// column values.
template <typename T, int QP, typename Z, typename SUF = NoSuffix>
__device__ __forceinline__ T dtw_band(int m, int w, int seg_len, int nseg, T ub, const Z &z, int &t_stop,
                                      const SUF &suf = SUF()) {
    if constexpr (Z::kFused) {
        return dtw_band_fused<T, QP, Z, SUF>(m, w, seg_len, nseg, ub, z, t_stop, suf);
    } else {
    const int lane = (int)(threadIdx.x & 63);
    const T INF = dev_inf<T>();
    const int seg = lane / seg_len;
    const int x = lane - seg * seg_len;
    const bool valid = seg < nseg;
    // lanes of this lane's segment
    const unsigned long long segmask =
        valid ? (seg_len >= 64 ? ~0ull : ((1ull << seg_len) - 1ull) << (seg * seg_len)) : 0ull;
    T E[QP], O[QP];
    int e_off[QP];
    bool inE[QP], inO[QP];
    // the slot of offset 0 holds D(-1, -1) = 0 before the first step
    const int idx0 = w >> 1;              // pair index of offset 0
    const bool zero_even = (w & 1) == 0;  // offset 0 is the even member of its pair
#pragma unroll
    for (int k = 0; k < QP; ++k) {
        const int xk = x * QP + k;
        e_off[k] = 2 * xk - w;
        inE[k] = valid && e_off[k] <= w;
        inO[k] = valid && e_off[k] + 1 <= w;
        E[k] = (valid && xk == idx0 && zero_even) ? (T)0 : INF;
        O[k] = (valid && xk == idx0 && !zero_even) ? (T)0 : INF;
    }
    const int mm1 = m - 1;
    const int t_end = 2 * m - 2;
    auto clampi = [mm1](int v) {
        if constexpr (Z::kPadded)
            return v;
        else
            return min(max(v, 0), mm1);
    };
    // offsets used for the z-value indices: lanes past the band (their cells are forced to
    // +inf) read at offset w + 1, so that the indices stay within the padded range
    int e_ix[QP];
#pragma unroll
    for (int k = 0; k < QP; ++k) e_ix[k] = min(e_off[k], w + 1);
    auto odd_step = [&](int r[QP], T zr[QP]) __attribute__((always_inline)) {
        // odd offsets: up = next even (t-1), left = own even (t-1), diag = own odd (t-2)
        const T right_in = from_right<T>(INF, E[0]);
#pragma unroll
        for (int k = 0; k < QP; ++k) {
            const T up = k + 1 < QP ? E[k + 1] : right_in;
            const T d = zr[k] - z.col(clampi(r[k] + e_ix[k] + 1));
            const T v = fma(d, d, min3(O[k], up, E[k]));
            O[k] = inO[k] ? v : INF;
        }
    };
    bool alive = valid;
    t_stop = t_end;
    int t = 0;
    int r[QP];
    T zr[QP];
    if (!zero_even) {
        // t = 0 has odd parity: only the odd offsets, on row (0 - e - 1) / 2
#pragma unroll
        for (int k = 0; k < QP; ++k) {
            r[k] = (0 - e_ix[k] - 1) >> 1;
            zr[k] = z.row(clampi(r[k]));
        }
        odd_step(r, zr);
        t = 1;
    }
    // z-values are read one iteration ahead (their latency overlaps a whole iteration): the
    // row value of the pair of steps, and the two columns r + e and r + e + 1; the odd step's
    // column is the next iteration's even-step column.
    T zce[QP], zco[QP];
#pragma unroll
    for (int k = 0; k < QP; ++k) {
        r[k] = (t - e_ix[k]) >> 1;
        zr[k] = z.row(clampi(r[k]));
        zce[k] = z.col(clampi(r[k] + e_ix[k]));
        zco[k] = z.col(clampi(r[k] + e_ix[k] + 1));
    }
    // One pair of steps: the even offsets at t and, when `odd`, the odd offsets at t + 1.
    auto step_pair = [&](bool odd) __attribute__((always_inline)) {
        T zr_n[QP], zco_n[QP];
#pragma unroll
        for (int k = 0; k < QP; ++k) {
            zr_n[k] = z.row(clampi(r[k] + 1));
            zco_n[k] = z.col(clampi(r[k] + e_ix[k] + 2));
        }
        // even offsets at t: up = own odd (t-1), left = previous odd (t-1), diag = own even (t-2)
        const T left_in = from_left<T>(INF, O[QP - 1]);
#pragma unroll
        for (int k = 0; k < QP; ++k) {
            const T left = k > 0 ? O[k - 1] : left_in;
            const T d = zr[k] - zce[k];
            const T v = fma(d, d, min3(E[k], O[k], left));
            E[k] = inE[k] ? v : INF;
        }
        if (odd) {
            // odd offsets: up = next even (t), left = own even (t), diag = own odd (t-1)
            const T right_in = from_right<T>(INF, E[0]);
#pragma unroll
            for (int k = 0; k < QP; ++k) {
                const T up = k + 1 < QP ? E[k + 1] : right_in;
                const T d = zr[k] - zco[k];
                const T v = fma(d, d, min3(O[k], up, E[k]));
                O[k] = inO[k] ? v : INF;
            }
        }
#pragma unroll
        for (int k = 0; k < QP; ++k) {
            ++r[k];
            zr[k] = zr_n[k];
            zce[k] = zco[k];
            zco[k] = zco_n[k];
        }
    };
    // Abandon test after the pair of steps at tc (tc = 14 or 15 mod 16); true when no segment
    // of the wave is left.  The segment is over ub when its minimum over both live
    // anti-diagonals is; both tests are monotone in the value, so that holds exactly when every
    // lane of the segment is over on its own cells: one ballot instead of a cross-lane minimum.
    auto check = [&](int tc) __attribute__((always_inline)) {
        T v = INF;
#pragma unroll
        for (int k = 0; k < QP; ++k) v = min3(v, E[k], O[k]);
        bool over = v > ub;
        if constexpr (SUF::enabled) {
            const T sb = (T)suf.at(tc);  // every lane takes part (the functor may shuffle)
            over = over || (v + sb) * (T)(1.0 - 0x1p-9) > ub;
        }
        const unsigned long long keep = __builtin_amdgcn_ballot_w64(!over);
        if (alive && (keep & segmask) == 0ull) {
            alive = false;
            t_stop = tc + 1;
        }
        return __builtin_amdgcn_ballot_w64(alive) == 0ull;
    };
    // Blocks of 8 pairs of steps (t = t0 + 16 k .. + 15, t0 = t mod 2 here) end in a test, as
    // long as the block's last odd step t + 15 comes before the last anti-diagonal; in the
    // remaining tail (t + 15 >= t_end) no test falls, since a test needs tc + 1 < t_end.
    bool done = false;
    while (t + 15 < t_end) {
#pragma unroll
        for (int u = 0; u < 8; ++u) step_pair(true);
        t += 16;
        if (check(t - 2)) {
            done = true;
            break;
        }
    }
    if (!done)
        for (; t <= t_end; t += 2) step_pair(t + 1 <= t_end);
    // result: offset 0 of the lane's segment
    const int src_lane = valid ? seg * seg_len + idx0 / QP : lane;
    const int k0 = idx0 % QP;
    T mine = INF;
#pragma unroll
    for (int k = 0; k < QP; ++k)
        if (k == k0) mine = zero_even ? E[k] : O[k];
    const T res = __shfl(mine, src_lane, 64);
    if (!alive || !valid) return INF;
    if (res > ub) {
        t_stop = t_end;
        return INF;
    }
    return res;
    }
}

}  // namespace dtwj
