// The fused top-1 join on gfx950: a persistent diagonal sweep over the pair triangles of one
// or many series with an in-wave pruning cascade (reference: the host tile loop of
// GPU_version1/src/new_dtw_motifGUI_malloc.cu:622-740 driving GLOBAL_DIAG,
// GPU_version1/lib/GLOBAL_DIAG.cu:29-73, CUB compaction and the per-pair LB_Keogh + DTW kernel,
// GPU_version1/src/dtw_kernal.cu:174-267).
//
// Work item = (series s, diagonal group g, row segment): 256 consecutive diagonal offsets
// d = d0 + 4*lane + q (q = 0..3, four diagonals per lane so the per-column data is reused
// from registers) over rows_per_item consecutive rows i (pair (i, j = i + d)).  The items of
// all series form one list claimed from a device counter (the device work queue of a batch).
// Each series has its own best-so-far, a packed 64-bit word (cost bits | i) read at every LDS
// chunk and every cascade batch and updated with atomicMin, so every wave prunes with the best
// value found anywhere on the device.  No host synchronisation inside the join.
//
// Stages per pair (each a valid lower bound of the squared banded DTW cost):
//  1. two-sided ball bound (all pairs; cf. diag_fast, GPU_version1/lib/diag_fast.cu:13-90):
//     the envelope box of window i lies in the ball with centre M_i = (U_i+L_i)/2 and radius
//     R_i = ||U_i-L_i||/2 (normalised), so LB_Keogh(z_j, env_i) >= (||z_j - M_i|| - R_i)^2.
//     ||z_j - M_i||^2 = m - 2 X/(sigma_i sigma_j) + ||M_i||^2, X = sum (a_j - mu_j)(v_i - muv_i),
//     updated in O(1) per step with the MPX identity
//     X(i+1,j+1) = X(i,j) + dfa[j] dgv[i] + dfv[i] dga[j], re-anchored exactly (FP64) at the
//     start of every item; a per-item rounding margin keeps the FP32 test conservative.
//     Per pair: 4 FMA updates + 2 FMA tests + 1 min; the column data of a lane's four
//     diagonals rotates through registers (one LDS read per table and 4 rows).  Where the
//     two-sided test leaves almost nothing (narrow bands), a wave sweeps its next items with
//     the one-sided test (X^A only: 2 FMA updates + 1 FMA test per pair, three tables) and
//     returns to two-sided when survivors grow; the choice is made per item from the
//     previous item's survivor rate.
//  2. LB_Kim over the 3x3 corners at both ends (cf. CPU_version1/lb_kim.cpp:10-73).
//  3. two-sided LB_Keogh against the raw-series envelope normalised by the other window's
//     statistics (cf. lb_keogh_stride_with_nomalise, GPU_version1/src/LB_keogh_warp.cu:74-96),
//     early-abandoned every 16 positions; a lane whose pair is decided takes the next queued
//     pair, so lanes stay busy whatever the abandon points.
//  4. banded DTW on the wavefront (dtw_device.hpp), early-abandoned against the best-so-far.
// Ball survivors are appended to a per-wave queue in global memory, one 4-byte entry per lane
// and row (row | lane | 4-bit complement of the mask of the lane's surviving diagonals: the
// sign bits of the four test values).  The queue is drained
// completely at the end of every 64-row chunk (earlier only if it could overflow): the chunk's
// series / envelope windows and window statistics are then staged in LDS over the sweep's
// column tables, the entries are expanded into a pair list, and stages 2-4 read LDS, not
// scattered global memory, one pair per lane.
#include <type_traits>

#include "dtw_device.hpp"
#include "fp16_env.hpp"
#include "../../include/dtwj.h"
#include "internal.hpp"

namespace dtwj {

namespace {

constexpr int KD = 4;                  // diagonals per lane
constexpr int DG = 64 * KD;            // diagonals per item
constexpr int RC = 64;                 // rows per cascade part (the drain stages RC rows at a time)
constexpr int JW = RC + DG + 8;        // columns an RC-row part spans
constexpr int NJARR = 6;               // column arrays: dfa dga dfv dgv 1/sigma thetaB
constexpr int RC2 = 128;               // rows per chunk of the two-sided sweep (6 + 6 tables)
constexpr int JW2 = RC2 + DG + 8;      // its staged column entries
constexpr int kSweepFloats = NJARR * JW2 + NJARR * RC2;  // column + row tables
constexpr int RC1 = 256;               // rows per chunk of the one-sided sweep (3 + 3 tables)
constexpr int JW1 = RC1 + DG + 8;
static_assert(3 * JW1 + 3 * RC1 <= kSweepFloats, "one-sided tables fit the two-sided area");
static_assert(RC1 % RC == 0 && RC2 % RC == 0, "the cascade drains a chunk in RC-row parts");
constexpr int PLCAP = 64 * KD;         // expanded pairs of one 64-entry queue batch
constexpr int Q2CAP = 256;             // LB_Keogh queue (packed pairs); it runs from Q2CAP - 64 on
constexpr int Q3CAP = 128;             // DTW queue (packed pairs)
constexpr int QPCAP = 128;             // segment-mean queue (packed pairs): < 64 left + one LB_Kim batch
constexpr int QLIM = kJoinQueueCap - 8 * 64;  // drain before an 8-row group could overflow

enum { C_SWEPT = 0, C_BALL, C_KIM, C_KEOGH, C_DTW_AB, C_CELLS, C_KEOGH_ELEMS, C_SWEPT_ONE, C_PAA, C_NCOUNTERS };
static_assert(C_NCOUNTERS == 9, "the host reads 9 counters per series");
constexpr int kWcnt = 10;  // LDS stage counters (80 bytes keep the dynamic area 16-byte aligned)
static_assert(kWcnt >= C_NCOUNTERS && kWcnt % 2 == 0, "stage counters");

// LDS words of the cascade staging for window length m (overlays the column tables): series
// and envelope of the chunk's row windows and column windows (upper and lower envelope in FP32,
// or, with pw, one word of packed FP16 widths), and the centred mean and 1/sigma of those
// windows.
__host__ __device__ constexpr int cascade_floats(int m, bool pw, bool gs = false) {
    return (pw ? 2 : 3) * (RC + m) + (pw ? 2 : 3) * (RC + DG + m) + 2 * RC + (gs ? 0 : 2 * (RC + DG));
}

__host__ __device__ constexpr int union_floats(int m, bool pw, bool gs = false) {
    return cascade_floats(m, pw, gs) > kSweepFloats ? cascade_floats(m, pw, gs) : kSweepFloats;
}

// gs: the join reads the mean and 1/sigma of the cascade's column windows from global memory
// instead of staging them in LDS, its DTW queue holds 64 pairs and its LB_Keogh queue 128 (it
// runs from 64 queued pairs on).  Launches of two offset pairs per lane take that form where
// it fits more waves per CU (their staging of long windows sets the count: at m = 1024,
// 22.3 KiB per wave, 7 waves instead of 6); the DTW runs one pair per wave there, so a deeper
// queue buys nothing.
__host__ __device__ constexpr int dtw_queue_cap(bool gs) { return gs ? 64 : Q3CAP; }
__host__ __device__ constexpr int keogh_queue_cap(bool gs) { return gs ? 128 : Q2CAP; }

// Join kernels whose cascade stages packed envelope widths: those of bands of 129..255 offsets
// (two offset pairs per lane), which come with long windows and run two waves per SIMD.
__host__ __device__ constexpr bool packed_envelope(int Q) { return Q == 2; }

struct Thr {
    float T2;      // squared-cost threshold (best-so-far inflated by the FP32 DTW error)
    float T;       // sqrt(T2), for the ball bound
    float kim;     // LB_Kim threshold
    float dtw_ub;  // DTW abandon / candidate threshold
};

__device__ inline float dtw_err(float D, const JoinArgs &A) {
    const float e = A.eps_z;
    return A.rel_dtw * D + 4.0f * e * sqrtf(2.0f * A.m * D) + 8.0f * A.m * e * e;
}

__device__ inline unsigned long long load_bsf(const JoinArgs &A) {
    return __hip_atomic_load(A.bsf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Thresholds of every stage from a best-so-far word (an older word is larger: still valid).
// The word's cost is at most that of a pair the join has seen as an exact (FP64) DTW cost
// (push_candidate records that bound, not the FP32 cost), so it serves as T2 as it is.
__device__ inline Thr threshold_of(unsigned long long key, const JoinArgs &A) {
    const float cost = bsf_cost(key);
    Thr t;
    t.T2 = cost;
    t.T = sqrtf(t.T2) * (1.0f + 0x1p-20f);
    const float e = A.eps_z;
    t.kim = t.T2 + 4.0f * e * sqrtf(12.0f * t.T2) + 48.0f * e * e + 0x1p-20f * t.T2;
    t.dtw_ub = t.T2 + dtw_err(t.T2, A);
    return t;
}

__device__ inline Thr load_threshold(const JoinArgs &A) { return threshold_of(load_bsf(A), A); }

// Bound on the error of one FP32 cell difference of the pair of windows with FP32 centred means
// mi, mj and 1/sigma ii, ij, in each form the kernels use (StagedZ, StagedZPlain, the seed DTW's
// (x - mean) (1/sigma)): each rounds the centred sample, the mean and 1/sigma once and takes at
// most five roundings more, so with |z| <= sqrt(m) inside a window
// e <= 2^-24 (max|x| (ii + ij) + 4 (|mi| ii + |mj| ij) + 8 sqrt(m)); the 2^-20 covers the
// second-order terms and this evaluation.  c8m = 8 sqrt(m).  With the pair's own 1/sigma and
// means in place of the series maxima it is far below the 2 eps_z of the series-wide model.
__device__ inline float pair_cell_err(const JoinArgs &A, float c8m, float mi, float ii, float mj, float ij) {
    return 0x1p-24f * (1.f + 0x1p-20f) * fmaf(A.amax, ii + ij, fmaf(4.f, fabsf(mi) * ii + fabsf(mj) * ij, c8m));
}

// (1 + rel_dtw) (r + e sqrt(2m))^2 for a pair with cell-difference error bound e (s2m =
// sqrt(2m)).  By Minkowski's inequality over the at most 2m cells of a path and the FP32
// accumulation of a path sum (at most 2m u relative, inside rel_dtw, which also covers the
// rounding of r):
//  * r = sqrt(C): a pair whose exact cost is at most C has an FP32 DTW of at most this (the
//    FP32 DTW is at most its FP32 sum along the exact optimal path);
//  * r = sqrt(FP32 DTW cost): the pair's exact cost is at most this (the exact cost is at most
//    the exact sum along the path the FP32 DTW took).
__device__ inline float widen_cost(float r, float e, float s2m, const JoinArgs &A) {
    const float q = fmaf(e, s2m, r);
    return (1.f + A.rel_dtw) * (q * q);
}

// DTW threshold of one pair: a pair whose exact cost is at most T2 (sT2 = sqrt(T2)) stays at or
// below it; the series-wide dtw_ub is used where smaller.
__device__ inline float pair_dtw_ub(const Thr &t, float sT2, float e, float s2m, const JoinArgs &A) {
    return fminf(widen_cost(sT2, e, s2m, A), t.dtw_ub);
}

// Records a pair whose FP32 cost is within the error margin of the best-so-far (re-scored in
// FP64 after the join) and lowers the best-so-far to key_cost, a bound on the pair's exact cost
// (widen_cost of its FP32 cost).  The push counter is 64-bit, so it keeps counting past
// cand_cap without wrapping: the host sees how many slots a re-run needs.
__device__ inline void push_candidate(const JoinArgs &A, int i, int j, float cost, float key_cost) {
    const unsigned long long slot = atomicAdd(A.n_cands, 1ull);
    if (slot < (unsigned long long)A.cand_cap) A.cands[slot] = make_int4(i, j, __float_as_int(cost), 0);
    atomicMin(A.bsf, pack_bsf(key_cost, (uint32_t)i));
}

__device__ inline float sq(float x) { return x * x; }
// x, with +0 replaced by -0 (equal as numbers; only the sign of a zero changes)
__device__ inline float no_pos_zero(float x) { return x == 0.0f ? -0.0f : x; }

__device__ inline void decode_item(const int64_t *item_start, int groups, int64_t item, int &g, int64_t &s) {
    int lo = 0, hi = groups - 1;  // g with item_start[g] <= item < item_start[g+1]
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (item_start[mid] <= item)
            lo = mid;
        else
            hi = mid - 1;
    }
    g = lo;
    s = item - item_start[g];
}

// Copies a uniform pointer into vector registers.  The join keeps many uniform values alive;
// pointers used only in a short loop are moved out of the scalar register file so that the
// compiler does not spill (and reload) scalar register tuples around that loop.
template <typename T>
using global_ptr = const __attribute__((address_space(1))) T *;
typedef float vf4 __attribute__((ext_vector_type(4)));  // 16-byte vector load
template <typename T>
__device__ inline global_ptr<T> in_vgprs(const T *p) {
    const uint64_t u = (uint64_t)p;
    uint32_t lo = (uint32_t)u, hi = (uint32_t)(u >> 32);
    asm volatile("v_mov_b32 %0, %1" : "=v"(lo) : "s"(lo));
    asm volatile("v_mov_b32 %0, %1" : "=v"(hi) : "s"(hi));
    return (global_ptr<T>)(((uint64_t)hi << 32) | lo);
}

__device__ inline unsigned int mbcnt64(unsigned long long msk) {
    return __builtin_amdgcn_mbcnt_hi((unsigned int)(msk >> 32), __builtin_amdgcn_mbcnt_lo((unsigned int)msk, 0u));
}

// Maximum of blkmax[which] over the 256-window blocks covering windows [lo, hi].
__device__ inline float range_max(const JoinArgs &A, int which, int64_t lo, int64_t hi) {
    float r = 0.f;
    for (int64_t b = lo >> 8; b <= (hi >> 8); ++b) r = fmaxf(r, A.blkmax[(int64_t)which * A.nblk + b]);
    return r;
}

// A DTW pair read from the staged chunk as raw values: the difference of a cell is formed as
// fma(x_c, nij, fma(x_r, ii, c)) with ii = 1/sigma_i, nij = -1/sigma_j and c = the difference
// of the two windows' -mean/sigma (pair_bias), i.e. z_i(r) - z_j(c) without normalising the
// columns.  Its error stays within the 2 eps_z the DTW error model allows per difference:
// 2^-24 (6 max|x| max(1/sigma) + 8 max|mean| max(1/sigma) + |d|) <= 2^-21 (max|x| +
// max|mean|) max(1/sigma) + 2^-23 sqrt(m).  Indices may leave the window by up to
// (w + 3) / 2 + 1 on either side: the cascade layout keeps finite staged values there (see
// CascadeLDS).
struct StagedZ {
    static constexpr bool kPadded = true;
    static constexpr bool kFused = true;
    const float *xi, *xj;
    float ii, nij, c;
    __device__ float xrow(int r) const { return xi[r]; }
    __device__ float xcol(int q) const { return xj[q]; }
};
// The same pair as z-values, each normalised on its own, x / sigma - mean / sigma in one FMA
// (nmi = -mean / sigma): the form of the kernels with one offset pair per lane (bands up to
// 63), where the fused form measured slower through the code generated around the sweep.
struct StagedZPlain {
    static constexpr bool kPadded = true;
    static constexpr bool kFused = false;
    const float *xi, *xj;
    float ii, nmi, ij, nmj;
    __device__ float row(int r) const { return fmaf(xi[r], ii, nmi); }
    __device__ float col(int c) const { return fmaf(xj[c], ij, nmj); }
};

// c = (-mu_i / sigma_i) - (-mu_j / sigma_j) of a staged pair, with the products rounded on
// their own: the DTW and the suffix bound of the same pair must obtain the same bits.
__device__ inline float pair_bias(float mi, float ii, float mj, float ij) {
    DTWJ_NO_CONTRACT
    const float nmi = -mi * ii, nmj = -mj * ij;
    return nmi - nmj;
}

// Suffix bound of the DTW of each segment's pair (see dtw_device.hpp): block kappa holds the
// LB_Keogh terms (both directions) of rows / columns [8 (kappa + 1) + ceil(w/2), + 8), and the
// check at anti-diagonal t (t = 16 k + 14 or + 15) needs the sum over the blocks kappa >= k = t / 16.
// Lane x of a segment holds the blocks kappa = x + seg_len q, q < SUF_Q, as reverse prefix sums
// max(sum A, sum B) over all blocks from kappa on.
constexpr int SUF_Q = 4;
struct LaneSuffix {
    static constexpr bool enabled = true;
    float reg[SUF_Q];
    int base, seg_len, nq;  // lane of the segment's first block, segment length, rounds used
    __device__ float at(int t) const {
        const int k = t >> 4;
        const int q = k / seg_len;
        float v = 0.f;
#pragma unroll
        for (int i = 0; i < SUF_Q; ++i)
            if (i == q) v = reg[i];
        const float r = __shfl(v, base + (k - q * seg_len), 64);
        return q < nq ? r : 0.f;
    }
};

// LDS layout of the cascade's staging (valid while draining) for window length m: mean and
// LDS layout of the cascade's staging (valid while draining) for window length m: mean and
// 1/sigma of the part's RC rows and of its RC + DG columns, then the windows of the rows (CI
// samples) and of the columns (CJ samples).  Without PW: series, upper and lower envelope of
// the rows, then the same of the columns, all FP32.  With PW: the two series, then the envelope
// widths of the same windows, two FP16 values per word (lo: U - a, hi: a - L, both rounded up,
// so a + lo >= U and a - hi <= L in FP32: a widened envelope, under which LB_Keogh and the DTW
// suffix bound stay lower bounds) in 2/3 of the space.  The DTW reads up to
// (w + 3) / 2 + 1 <= 2 RC + 2 (RC + DG) values before and after a series window without
// clamping; those reads land on finite values: before the row series the statistics, after it
// the row upper envelope (PW: the column series), before the column series the row lower
// envelope (PW: the row series), after it the column upper envelope (PW: the row widths, whose
// words read as floats are finite: sign 0 and exponent field <= 248, both halves being
// non-negative FP16 values up to +inf).  Built where it is used from an opaque copy of m, so that
// the addresses are not kept live (in scalar registers that would spill) through the sweep.
template <bool PW>
struct CascadeLDS {
    float *sai, *saj, *rmu, *ris, *cmu, *cis;
    float *sui = nullptr, *sli = nullptr, *suj = nullptr, *slj = nullptr;  // without PW
    uint32_t *swi = nullptr, *swj = nullptr;                              // with PW
    // gs: the column statistics are not staged (cmu, cis unused)
    __device__ CascadeLDS(float *base, int m, bool gs = false) {
        asm volatile("" : "+s"(m));
        const int CI = RC + m, CJ = RC + DG + m;
        rmu = base;
        ris = rmu + RC;
        if (gs) {
            cmu = cis = nullptr;
            sai = ris + RC;
        } else {
            cmu = ris + RC;
            cis = cmu + (RC + DG);
            sai = cis + (RC + DG);
        }
        if constexpr (PW) {
            saj = sai + CI;
            swi = reinterpret_cast<uint32_t *>(saj + CJ);
            swj = swi + CI;
        } else {
            sui = sai + CI;
            sli = sui + CI;
            saj = sli + CI;
            suj = saj + CJ;
            slj = suj + CJ;
        }
    }
    // widened envelope (upper, lower) at a sample x with packed widths wb
    __device__ static void env(float x, uint32_t wb, float &up, float &lo) { fp16_env_decode(x, wb, up, lo); }
    // envelope of the row window at position k (series value x = sai[k]), of the column window
    __device__ void env_i(int k, float x, float &up, float &lo) const {
        if constexpr (PW)
            env(x, swi[k], up, lo);
        else {
            up = sui[k];
            lo = sli[k];
        }
    }
    __device__ void env_j(int k, float x, float &up, float &lo) const {
        if constexpr (PW)
            env(x, swj[k], up, lo);
        else {
            up = suj[k];
            lo = slj[k];
        }
    }
};
static_assert((511 + 3) / 2 + 1 <= 2 * RC + 2 * (RC + DG),
              "DTW reads before a row window stay in the statistics for the widest band, w = 511");
static_assert((127 + 3) / 2 + 1 <= 2 * RC,
              "without staged column statistics (bands up to w = 127), in the row statistics");

// Suffix bound of a DTW from the pair's LB_Keogh totals, which LB_Keogh formed term by term with
// the DTW's own arithmetic (fz_term): the sum over the rows (columns) from R on is the total minus
// the prefix below R, and the prefix is formed only as far as the checks reach, one chunk of rows
// at a time.  One pair per wave (G = false): 64-row chunks, lane x forms the terms of row
// 64 h + x.  Several pairs per wave (G): the pair of DTW segment s has the lane group
// [s slen, (s + 1) slen), slen = 64 / (pairs per wave), whose lane x forms the terms of row
// slen h + x of that pair, and a segmented prefix sum runs over each group.  An FP32 sum of n
// non-negative terms lies within (1 +- 1.01 n u) of the exact sum (u = 2^-24), so total
// (1 - 2^-10) - prefix (1 + 2^-10) is at most the exact sum of the suffix's terms for
// m <= 16000; each term bounds the FP32 costs of its row's (column's) cells (see StagedZ), as
// the DTW's check requires.
template <bool PW, bool G>
struct TotSuffix {
    static constexpr bool enabled = true;
    const CascadeLDS<PW> *L;
    int si, sj, m, w2;    // staged windows of the lane group's pair, window length, ceil(w / 2)
    float ii, nij, c;     // that pair's StagedZ coefficients
    float totC, totR;     // LB_Keogh totals of the lane's DTW pair: columns of j against env_i,
                          // rows of i against env_j
    int h;                // last chunk of the prefix formed (-1: none)
    float preC, preR;     // inclusive prefix sums up to row slen h + x of the group's pair
    int slen, gx, gbase, rbase;  // (G) lanes per group, index in the group, the group's first
                                 // lane, first lane of the DTW pair's group
    bool gvalid;                 // (G) the lane's group has a pair
    __device__ float at(int t) {
        const int R = 8 * ((t >> 4) + 1) + w2;  // first row of the suffix at this check
        if (R >= m) return 0.f;
        const int q = R - 1;
        const int lane = (int)(threadIdx.x & 63);
        const int sl = G ? slen : 64;
        const int x = G ? gx : lane;
        const int last = G ? min(gbase + sl - 1, 63) : 63;
        while (G ? (h + 1) * sl <= q : h < (q >> 6)) {
            const float cC = h >= 0 ? __shfl(preC, last, 64) : 0.f;
            const float cR = h >= 0 ? __shfl(preR, last, 64) : 0.f;
            ++h;
            const int r = sl * h + x;
            float sC = 0.f, sR = 0.f;
            if ((!G || gvalid) && r < m) {
                const float xi = L->sai[si + r], xj = L->saj[sj + r];
                float Ui, Li, Uj, Lj;
                L->env_i(si + r, xi, Ui, Li);
                L->env_j(sj + r, xj, Uj, Lj);
                const float dlo = fmaf(xj, nij, fmaf(Li, ii, c)), dhi = fmaf(xj, nij, fmaf(Ui, ii, c));
                const float ta = fmaxf(fmaxf(dlo, -dhi), 0.f);
                const float zr = fmaf(xi, ii, c);
                const float dU = fmaf(Uj, nij, zr), dL = fmaf(Lj, nij, zr);
                const float tb = fmaxf(fmaxf(dU, -dL), 0.f);
                sC = ta * ta;
                sR = tb * tb;
            }
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const float yC = __shfl_up(sC, off, 64), yR = __shfl_up(sR, off, 64);
                if (x >= off) {
                    sC += yC;
                    sR += yR;
                }
            }
            preC = sC + cC;
            preR = sR + cR;
        }
        const int src = G ? rbase + (q - h * sl) : (q & 63);
        const float pC = __shfl(preC, src, 64), pR = __shfl(preR, src, 64);
        const float vC = totC * (1.f - 0x1p-10f) - pC * (1.f + 0x1p-10f);
        const float vR = totR * (1.f - 0x1p-10f) - pR * (1.f + 0x1p-10f);
        return fmaxf(fmaxf(vC, vR), 0.f);
    }
};

// Segment-mean bound of the window pair (gi, gj): len sum_s max(Z_j - U_i, L_i - Z_j, 0)^2 and
// the same with i and j swapped, the larger of the two (rows of the series' table, see kPaaSeg).
// For the LB_Keogh terms t_k of a segment's samples, sum t_k^2 >= (sum t_k)^2 / len
// (Cauchy-Schwarz) and sum t_k >= len max(Z - U, L - Z, 0), so the bound is at most LB_Keogh
// against the raw-series envelope, itself at most the banded DTW.  In FP32 each term costs one
// rounded subtraction and one FMA: the result is at most 20 * 2^-24 above the bound of the
// table values, which a caller's (1 - 2^-17) scaling covers.
__device__ inline float paa_lb(const float *paa, int64_t gi, int64_t gj, int nseg, float len) {
    const float4 *ti = reinterpret_cast<const float4 *>(paa + gi * 3 * nseg);
    const float4 *tj = reinterpret_cast<const float4 *>(paa + gj * 3 * nseg);
    float accA = 0.f, accB = 0.f;
    auto term = [](float z, float u, float l, float &acc) __attribute__((always_inline)) {
        const float t = fmaxf(fmaxf(z - u, l - z), 0.f);
        acc = fmaf(t, t, acc);
    };
#pragma unroll
    for (int q = 0; q < nseg / 4; ++q) {
        const float4 ui = ti[q], li = ti[nseg / 4 + q], zi = ti[nseg / 2 + q];
        const float4 uj = tj[q], lj = tj[nseg / 4 + q], zj = tj[nseg / 2 + q];
        term(zj.x, ui.x, li.x, accA);
        term(zj.y, ui.y, li.y, accA);
        term(zj.z, ui.z, li.z, accA);
        term(zj.w, ui.w, li.w, accA);
        term(zi.x, uj.x, lj.x, accB);
        term(zi.y, uj.y, lj.y, accB);
        term(zi.z, uj.z, lj.z, accB);
        term(zi.w, uj.w, lj.w, accB);
    }
    return len * fmaxf(accA, accB);
}
static_assert(paa_segments(8) % 8 == 0 && paa_segments(1 << 20) % 8 == 0 && kPaaSegMax >= paa_segments(1 << 20),
              "the table is read as float4 vectors, 8 segments per round");

// The same bound as a test against thr (true: the pair survives), with early exit: direction A
// (window j's means against window i's envelope means), then B, each in rounds of 8 segments;
// a lane stops loading as soon as its partial sum (a lower bound of the full one)
// exceeds thr, so the table rows of pruned pairs are mostly not read.  Exits exactly when
// paa_lb(...) (1 - 2^-17) > thr would, up to the rounding of the partial sums.
__device__ inline bool paa_pass(const float *paa, int64_t gi, int64_t gj, int nseg, float len, float thr) {
    const float4 *ti = reinterpret_cast<const float4 *>(paa + gi * 3 * nseg);
    const float4 *tj = reinterpret_cast<const float4 *>(paa + gj * 3 * nseg);
    const float scale = len * (1.0f - 0x1p-17f);
    bool alive = true;
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
        const float4 *te = dir == 0 ? ti : tj;  // envelope side
        const float4 *tz = dir == 0 ? tj : ti;  // series side
        float acc = 0.f;
#pragma unroll
        for (int h = 0; h < kPaaSegMax / 8; ++h) {
            if (alive && 8 * h < nseg) {
                float4 u[2], l[2], z[2];
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    u[k] = te[2 * h + k];
                    l[k] = te[nseg / 4 + 2 * h + k];
                    z[k] = tz[nseg / 2 + 2 * h + k];
                }
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const float zz[4] = {z[k].x, z[k].y, z[k].z, z[k].w};
                    const float uu[4] = {u[k].x, u[k].y, u[k].z, u[k].w};
                    const float ll[4] = {l[k].x, l[k].y, l[k].z, l[k].w};
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const float t = fmaxf(fmaxf(zz[c] - uu[c], ll[c] - zz[c]), 0.f);
                        acc = fmaf(t, t, acc);
                    }
                }
                alive = !(acc * scale > thr);
            }
        }
    }
    return alive;
}

// Packed (row, column) offsets of a pair inside the staged chunk (both < 2^16).
__device__ inline uint32_t pack_pair(int li, int lj) { return (uint32_t)li | ((uint32_t)lj << 16); }
__device__ inline int pair_row(uint32_t p) { return (int)(p & 0xffffu); }
__device__ inline int pair_col(uint32_t p) { return (int)(p >> 16); }

// series: nseries join descriptors; item_base: nseries + 1 cumulative item counts (the
// device work queue is the range [0, item_base[nseries]) claimed through work_counter), then
// the claim-order multiplier;
// uni: floats of the LDS union area, sized for the longest window of the launch.
template <int Q, bool PW, bool SW, bool GS = false>
__device__ __forceinline__ void join_body(const JoinArgs *__restrict__ series, const int64_t *__restrict__ item_base,
                                          int32_t nseries, int32_t uni, unsigned long long *work_counter, uint32_t *queue) {
    extern __shared__ __align__(16) float smem[];  // 16-byte aligned: the sweep reads float4 rows
    const int lane = (int)threadIdx.x;
    uint32_t *q1 = queue + (size_t)blockIdx.x * kJoinQueueCap;  // this wave's survivor queue

    // Stage counters live in LDS (updated by lane 0 on the rare paths) so that they occupy no
    // scalar registers in the sweep; LB_Keogh positions are counted per lane.  They are added
    // to the series' counters whenever the wave moves to another series, and at the end.
    __shared__ __align__(16) unsigned long long wcnt[kWcnt];
    if (lane < kWcnt) wcnt[lane] = 0ull;
    auto count = [&](int which, unsigned long long v) __attribute__((always_inline)) {
        if (lane == 0) wcnt[which] += v;
    };
    unsigned long long c_kelem = 0;  // per lane: LB_Keogh positions evaluated
    uint32_t c_cells = 0;            // per lane (segment leaders): DTW cells, flushed below 2^31

    int cur = -1;  // series of the current item
    bool two_sided = true;    // sweep test of the current item
    int prev_series = -1;     // series of the previous item (mode history)
    int64_t prev_rows = 0;    // rows of the previous item
    int64_t item_entries = 0; // survivor queue entries of the current item
    const JoinArgs *Ap = series;
    auto flush_counters = [&]() __attribute__((always_inline)) {
        if (cur < 0) return;
        __syncthreads();
        if (lane == 0) {
            for (int c = 0; c < C_NCOUNTERS; ++c)  // C_KEOGH_ELEMS is counted per lane below
                if (wcnt[c]) atomicAdd(&Ap->counters[c], wcnt[c]);
        }
        if (c_kelem) atomicAdd(&Ap->counters[C_KEOGH_ELEMS], c_kelem);
        if (c_cells) atomicAdd(&Ap->counters[C_CELLS], (unsigned long long)c_cells);
        c_kelem = 0;
        c_cells = 0;
        __syncthreads();
        if (lane < kWcnt) wcnt[lane] = 0ull;
        __syncthreads();
    };

    int q1n = 0, q2n = 0, q3n = 0, qpn = 0;
    int64_t i_beg = 0, d0 = 0, c0 = 0;
    int64_t cb = 0;  // first row of the RC rows staged for the cascade
    int m = 0;
    Thr th{};
    unsigned long long thk = ~0ull;  // best-so-far word th was computed from in the cascade
    float *jarr = smem;                // NJARR x JW2 column tables of the sweep
    float *rarr = jarr + NJARR * JW2;  // NJARR x RC2 row tables of the sweep
    // cascade staging (valid only while draining), laid out per series from its window length
    uint32_t *plist = reinterpret_cast<uint32_t *>(smem + uni);  // expanded pairs of a batch
    uint32_t *q2 = plist + PLCAP;                                // LB_Keogh queue
    constexpr int Q2C = keogh_queue_cap(GS), Q3C = dtw_queue_cap(GS);
    uint32_t *q3 = q2 + Q2C;                                     // DTW queue
    uint32_t *qp = q3 + Q3C;                                     // segment-mean queue
    // kernels with several offset pairs per lane: the LB_Keogh totals of each DTW queue entry
    // (columns of window j against the envelope of window i, rows of i against that of j)
    float *q3ta = reinterpret_cast<float *>(qp + QPCAP), *q3tb = q3ta + Q3C;
    // GS: centred mean and 1/sigma of the part's column windows read from global memory (flat
    // arrays, so that neighbouring lanes read neighbouring words), window index clamped to the
    // part's last window in the series
    global_ptr<float> gcmu = nullptr, gcis = nullptr;
    int gclim = 0;
    auto col_mu = [&](const CascadeLDS<PW> &L, int k) __attribute__((always_inline)) -> float {
        if constexpr (GS)
            return gcmu[k < gclim ? k : gclim];
        else
            return L.cmu[k];
    };
    auto col_is = [&](const CascadeLDS<PW> &L, int k) __attribute__((always_inline)) -> float {
        if constexpr (GS)
            return gcis[k < gclim ? k : gclim];
        else
            return L.cis[k];
    };
    // the cascade's DTW, LB_Keogh and LB_Kim form cell differences from raw values (see
    // StagedZ): in the kernels with several offset pairs per lane, and in the one-pair-per-lane
    // kernels of launches whose widest band is 4 or more (SW; bands of 16 and more there also
    // take the DTW suffix bound from the LB_Keogh totals)
    constexpr bool FZ = Q >= 2 || SW;
    // In the cascade of the kernels with the fused arithmetic (FZ), th is recomputed (three
    // square roots) only when the best-so-far word differs from thk, the word it was last
    // computed from there; a wave moving to another series invalidates it (no packed word has
    // an all-ones cost field).  The sweep's chunk staging recomputes th from the current word
    // without updating thk: words only fall, so when a later word equals thk, th is from that
    // word or a lower one, both valid.  (The other kernels recompute it each time: their code
    // measured faster that way.)
    auto set_th = [&](unsigned long long key) __attribute__((always_inline)) {
        if constexpr (FZ) {
            if (key != thk) {
                th = threshold_of(key, *Ap);
                thk = key;
            }
        } else {
            th = threshold_of(key, *Ap);
        }
    };

    // ---- stage 4: DTW of every queued pair.  Bands up to 63 run several pairs per wave
    // (segments of w + 1 lanes); wider bands one pair per wave with QP offset pairs per lane.
    auto process_q3 = [&]() __attribute__((always_inline)) {
        const JoinArgs &A = *Ap;
        const CascadeLDS<PW> L(smem, m, GS);
        const int w = A.w;
        const float c8m = 8.f * sqrtf((float)m), s2m = sqrtf(2.f * (float)m);
        const int seg_len = w <= 63 ? w + 1 : 64;
        const int per_wave = 64 / seg_len;
        const int seg = lane / seg_len;
        const bool leader = lane == seg * seg_len;
        // The suffix sums of pair s are formed by the lanes [s slen, (s + 1) slen), slen =
        // 64 / per_wave >= seg_len: the lanes a segment leaves idle in the DTW share the block
        // sums (w = 51: 64 lanes instead of 52, one round for m = 512 instead of two).
        const int slen = 64 / per_wave;
        const int sseg = lane / slen;  // pair whose block sums this lane forms
        // suffix-bound abandon for bands of 16 and more (narrower bands end their DTWs within a
        // few checks anyway) when the blocks fit SUF_Q rounds and m keeps the FP32 margin valid
        const int suf_blocks = (m - 8 - ((w + 1) >> 1) + 7) / 8;  // blocks with a row below m
        const int suf_rounds = suf_blocks > 0 ? (suf_blocks + slen - 1) / slen : 0;
        const bool use_suffix = w >= 16 && m <= 8192 && suf_rounds <= SUF_Q;
        // the first batch uses the current thresholds; each later one the best-so-far word
        // loaded one batch earlier (an older word is only larger: still valid)
        unsigned long long key = load_bsf(A);
        for (int e0 = 0; e0 < q3n; e0 += per_wave) {
            const int nseg = (q3n - e0) < per_wave ? (q3n - e0) : per_wave;
            if (e0 > 0) {
                set_th(key);
                key = load_bsf(A);
            }
            int li = 0, lj = 0;
            float ii = 0.f, nmi = 0.f, ij = 0.f, nmj = 0.f;
            if (seg < nseg) {
                const uint32_t pr = q3[e0 + seg];
                li = pair_row(pr);
                lj = pair_col(pr);
                ii = L.ris[li];
                nmi = -L.rmu[li] * ii;
                ij = col_is(L, lj);
                nmj = -col_mu(L, lj) * ij;
            }
            // the pair's own error bound and DTW threshold (the same on every lane of its segment)
            const float mcj = col_mu(L, lj);
            const float perr = pair_cell_err(A, c8m, L.rmu[li], ii, mcj, ij);
            const float ub = seg < nseg ? pair_dtw_ub(th, sqrtf(th.T2), perr, s2m, A) : th.dtw_ub;
            using ZT = std::conditional_t<FZ, StagedZ, StagedZPlain>;
            ZT z;
            if constexpr (FZ)
                z = StagedZ{L.sai + li, L.saj + lj, ii, -ij, seg < nseg ? pair_bias(L.rmu[li], ii, mcj, ij) : 0.f};
            else
                z = StagedZPlain{L.sai + li, L.saj + lj, ii, nmi, ij, nmj};
            int t_stop = 0;
            float res;
            // block sums of the LB_Keogh terms of the lane's blocks and their reverse prefix sums
            // over the pair's blocks (rounds q from last to first, carrying the totals); the
            // check of a DTW segment reads the sums of its pair from lanes seg slen + ...
            LaneSuffix suf;
            suf.base = seg * slen;
            suf.seg_len = slen;
            // (the kernels with several offset pairs per lane take TotSuffix below instead)
            suf.nq = (use_suffix && !FZ) ? suf_rounds : 0;
#pragma unroll
            for (int q = 0; q < SUF_Q; ++q) suf.reg[q] = 0.f;
            if (suf.nq > 0) {
                const int x = lane - sseg * slen;
                const int sbase = sseg * slen;
                int si = 0, sj = 0;
                float sii = 0.f, snmi = 0.f, sij = 0.f, snmj = 0.f;
                if (sseg < nseg) {
                    const uint32_t pr = q3[e0 + sseg];
                    si = pair_row(pr);
                    sj = pair_col(pr);
                    sii = L.ris[si];
                    snmi = -L.rmu[si] * sii;
                    sij = col_is(L, sj);
                    snmj = -col_mu(L, sj) * sij;
                }
                const StagedZPlain zs{L.sai + si, L.saj + sj, sii, snmi, sij, snmj};
                float sc = 0.f;  // the fused form's pair bias, as the DTW forms it
                if constexpr (FZ) {
                    if (sseg < nseg) sc = pair_bias(L.rmu[si], sii, col_mu(L, sj), sij);
                }
                float cA = 0.f, cB = 0.f;
#pragma unroll
                for (int q = SUF_Q - 1; q >= 0; --q) {
                    suf.reg[q] = 0.f;
                    if (q < suf.nq) {
                        float vA = 0.f, vB = 0.f;
                        const int r0 = 8 * (x + slen * q + 1) + ((w + 1) >> 1);
                        if (sseg < nseg && r0 < m) {
                            // the lane's 8 rows in an order skewed by the lane index: at each
                            // step the lanes' addresses differ by 9 (odd), so no bank conflicts
#pragma unroll
                            for (int rr = 0; rr < 8; ++rr) {
                                const int r = r0 + ((rr + x) & 7);
                                if (r < m) {
                                    // Terms by the DTW's own arithmetic (StagedZ, or the
                                    // z-values of StagedZPlain): a cell's FP32 difference is
                                    // monotone in x_c and in x_r, so with the envelope values in
                                    // place of x_c (row r, direction A) or of x_r (column r,
                                    // direction B) the results bound every cell of that row /
                                    // column.
                                    const float xi = L.sai[si + r], xj = L.saj[sj + r];
                                    float Ui, Li, Uj, Lj;
                                    L.env_i(si + r, xi, Ui, Li);
                                    L.env_j(sj + r, xj, Uj, Lj);
                                    float ta, tb;
                                    if constexpr (FZ) {
                                        const float zr = fmaf(xi, sii, sc);
                                        const float dU = fmaf(Uj, -sij, zr), dL = fmaf(Lj, -sij, zr);
                                        const float dlo = fmaf(xj, -sij, fmaf(Li, sii, sc));
                                        const float dhi = fmaf(xj, -sij, fmaf(Ui, sii, sc));
                                        ta = fmaxf(fmaxf(dU, -dL), 0.f);
                                        tb = fmaxf(fmaxf(dlo, -dhi), 0.f);
                                    } else {
                                        // z-values (StagedZPlain): one FMA each, monotone
                                        const float zi = zs.row(r), zj = zs.col(r);
                                        const float uj = fmaf(Uj, sij, snmj), lwj = fmaf(Lj, sij, snmj);
                                        const float ui = fmaf(Ui, sii, snmi), lwi = fmaf(Li, sii, snmi);
                                        ta = fmaxf(fmaxf(zi - uj, lwj - zi), 0.f);
                                        tb = fmaxf(fmaxf(zj - ui, lwi - zj), 0.f);
                                    }
                                    vA = fmaf(ta, ta, vA);
                                    vB = fmaf(tb, tb, vB);
                                }
                            }
                        }
#pragma unroll
                        for (int off = 1; off < 64; off <<= 1) {
                            const float yA = __shfl_down(vA, off, 64), yB = __shfl_down(vB, off, 64);
                            if (x + off < slen) {
                                vA += yA;
                                vB += yB;
                            }
                        }
                        vA += cA;
                        vB += cB;
                        cA = __shfl(vA, sbase, 64);
                        cB = __shfl(vB, sbase, 64);
                        suf.reg[q] = fmaxf(vA, vB);
                    }
                }
            }
            if constexpr (FZ) {
                // suffix bound from the LB_Keogh totals (the margins hold up to m = 8192) for
                // bands of 16 and more; narrower ones run without it
                if (Q >= 2 && w > 63 && m <= 8192) {
                    TotSuffix<PW, false> ts{&L, li, lj, m, (w + 1) >> 1, ii, -ij, z.c, q3ta[e0], q3tb[e0], -1, 0.f, 0.f};
                    res = dtw_band<float, Q>(m, w, 64, 1, ub, z, t_stop, ts);
                } else if (SW && w >= 16 && m <= 8192) {
                    TotSuffix<PW, true> ts;
                    ts.L = &L;
                    ts.m = m;
                    ts.w2 = (w + 1) >> 1;
                    ts.slen = slen;
                    ts.gbase = sseg * slen;
                    ts.gx = lane - ts.gbase;
                    ts.gvalid = sseg < nseg;
                    ts.si = ts.sj = 0;
                    ts.ii = ts.nij = ts.c = 0.f;
                    if (ts.gvalid) {
                        const uint32_t pr = q3[e0 + sseg];
                        ts.si = pair_row(pr);
                        ts.sj = pair_col(pr);
                        ts.ii = L.ris[ts.si];
                        const float gij = col_is(L, ts.sj);
                        ts.nij = -gij;
                        ts.c = pair_bias(L.rmu[ts.si], ts.ii, col_mu(L, ts.sj), gij);
                    }
                    ts.rbase = seg * slen;
                    ts.totC = seg < nseg ? q3ta[e0 + seg] : 0.f;
                    ts.totR = seg < nseg ? q3tb[e0 + seg] : 0.f;
                    ts.h = -1;
                    ts.preC = ts.preR = 0.f;
                    res = dtw_band<float, 1>(m, w, seg_len, nseg, ub, z, t_stop, ts);
                } else if (w > 63) {
                    res = dtw_band<float, Q>(m, w, 64, 1, ub, z, t_stop);
                } else {
                    res = dtw_band<float, 1>(m, w, seg_len, nseg, ub, z, t_stop);
                }
            } else {
                if (Q == 1 || w <= 63)  // Q = 1 launches hold bands up to 63 only (2w + 1 <= 128)
                    res = dtw_band<float, 1>(m, w, seg_len, nseg, ub, z, t_stop, suf);
                else
                    res = dtw_band<float, Q>(m, w, 64, 1, ub, z, t_stop, suf);
            }
            const bool lead = leader && seg < nseg;
            const bool ab = lead && !(res <= ub);
            if (lead) {
                const unsigned long long cells = (unsigned long long)dtw_cells_upto(m, w, t_stop);
                if (cells + c_cells >= (1ull << 31)) {
                    atomicAdd(&A.counters[C_CELLS], cells + c_cells);
                    c_cells = 0;
                } else {
                    c_cells += (uint32_t)cells;
                }
            }
            count(C_DTW_AB, __popcll(__builtin_amdgcn_ballot_w64(ab)));
            if (lead && !ab) {
                const int gi = (int)(cb + li), gj = (int)(cb + d0 + lj);
                push_candidate(A, gi, gj, res, widen_cost(sqrtf(res), perr, s2m, A));
            }
        }
        q3n = 0;
    };

    // ---- stage 3: two-sided LB_Keogh over the queued pairs, one pair per lane; a lane takes
    // the next pair as soon as its pair is abandoned or complete.  Consumes the whole queue.
    auto process_q2 = [&]() __attribute__((always_inline)) {
        const JoinArgs &A = *Ap;
        const CascadeLDS<PW> L(smem, m, GS);
        __syncthreads();  // q2 entries were written by other lanes
        // th: the drain batch's thresholds (at most one batch old)
        const float sm = sqrtf((float)m);
        float T2 = th.T2, sqmT2 = sqrtf(m * T2), sT2 = sqrtf(T2);  // updated after each DTW run
        // FZ: LB_Keogh by the DTW's own arithmetic (a term bounds, exactly, every FP32 cell cost
        // of its row / column, and the FP32 sum of the m terms of a direction is at most
        // (1 + 1.01 m u) times their exact sum while the DTW's FMA chain loses at most 2 m u:
        // together at most 1 + 3.1 m u, u = 2^-24), so a pair is dropped only when its FP32 DTW
        // would exceed the pair's DTW threshold (pair_dtw_ub)
        const float zfac = 1.f + 4.f * (float)m * 0x1p-24f;
        const float c8m = 8.f * sm, s2m = sqrtf(2.f * (float)m);
        int taken = 0;     // queue entries handed out (uniform)
        bool act = false;  // this lane holds a pair
        int li = 0, lj = 0, k = 0;
        // Windows of 256 and more are taken in blocks of 16 positions spread over the window
        // (stride 16 within superblocks of 256 positions, the superblocks interleaved), so that
        // a pair's bound grows with samples of the whole window and abandons sooner than in
        // window order; positions past the last whole superblock (and shorter windows) follow
        // in order.  kb, kc: block and superblock of the lane's next strided block.
        constexpr int KS = 16;
        const int nsup = m / (16 * KS);      // superblocks
        const int nstr = 16 * KS * nsup;     // positions taken strided
        int kb = 0, kc = 0;
        float accA = 0.f, accB = 0.f, thrA = 0.f, thrB = 0.f, rA = 0.f, bA = 0.f, rB = 0.f, bB = 0.f;
        // the DTW queue is run from one place (below) whenever it fills up and at the end, so
        // that the DTW code is inlined once
        bool more = true;
        while (more) {
        for (;;) {
            // hand out queued pairs to idle lanes
            const bool need = !act;
            const unsigned long long nk = __builtin_amdgcn_ballot_w64(need);
            const int idx = taken + (int)mbcnt64(nk);
            if (need && idx < q2n) {
                const uint32_t pr = q2[idx];
                li = pair_row(pr);
                lj = pair_col(pr);
                const float isi = L.ris[li], isj = col_is(L, lj);
                const float mci = L.rmu[li], mcj = col_mu(L, lj);
                if constexpr (FZ) {
                    // rA = 1/sigma_i, rB = -1/sigma_j, bA = the pair's bias (StagedZ)
                    rA = isi;
                    rB = -isj;
                    bA = pair_bias(mci, isi, mcj, isj);
                    thrA = thrB = pair_dtw_ub(th, sT2, pair_cell_err(A, c8m, mci, isi, mcj, isj), s2m, A) * zfac;
                } else {
                    // raw envelope of window i mapped into the units of window j, and the reverse
                    rA = isi / isj;  // sigma_j / sigma_i
                    bA = mcj - mci * rA;
                    rB = isj / isi;  // sigma_i / sigma_j
                    bB = mci - mcj * rB;
                    const float eA = 0x1p-22f * ((A.amax * (1.f + rA) + fabsf(bA)) * isj + sm);
                    const float eB = 0x1p-22f * ((A.amax * (1.f + rB) + fabsf(bB)) * isi + sm);
                    // sigma^2 from the hardware reciprocal (1 ulp): the 2^-20 inflation covers it
                    const float rsj = __builtin_amdgcn_rcpf(isj), rsi = __builtin_amdgcn_rcpf(isi);
                    const float sj2 = rsj * rsj, si2 = rsi * rsi;
                    thrA = (T2 + 2.f * eA * sqmT2 + m * eA * eA + m * 0x1p-22f * T2) * sj2 * (1.f + 0x1p-20f);
                    thrB = (T2 + 2.f * eB * sqmT2 + m * eB * eB + m * 0x1p-22f * T2) * si2 * (1.f + 0x1p-20f);
                }
                accA = accB = 0.f;
                k = 0;
                kb = kc = 0;
                act = true;
            }
            taken += __popcll(nk);
            if (taken > q2n) taken = q2n;
            if (__builtin_amdgcn_ballot_w64(act) == 0ull) {
                more = false;
                break;
            }
            if (act) {
                // x - (U rA + bA) = (x - bA) - U rA and (L rA + bA) - x = L rA - (x - bA): one
                // subtraction and two FMAs per direction (rounding error <= 2u (|x| + |U rA| +
                // |bA|), inside the 4u of eA / eB)
                // FZ terms of one position: column value xj against the envelope of window i
                // (accA) and row value xi against that of window j (accB), exactly as the DTW
                // suffix bound forms them
                auto fz_term = [&](float xj, float xi, float Ui, float Li, float Uj, float Lj)
                                   __attribute__((always_inline)) {
                    const float dlo = fmaf(xj, rB, fmaf(Li, rA, bA)), dhi = fmaf(xj, rB, fmaf(Ui, rA, bA));
                    const float ta = fmaxf(fmaxf(dlo, -dhi), 0.f);
                    accA = fmaf(ta, ta, accA);
                    const float zr = fmaf(xi, rA, bA);
                    const float dU = fmaf(Uj, rB, zr), dL = fmaf(Lj, rB, zr);
                    const float tb = fmaxf(fmaxf(dU, -dL), 0.f);
                    accB = fmaf(tb, tb, accB);
                };
                auto term = [&](int kk) __attribute__((always_inline)) {
                    const float xj = L.saj[lj + kk], xi = L.sai[li + kk];
                    float Ui, Li, Uj, Lj;
                    L.env_i(li + kk, xi, Ui, Li);
                    L.env_j(lj + kk, xj, Uj, Lj);
                    if constexpr (FZ) {
                        fz_term(xj, xi, Ui, Li, Uj, Lj);
                        return;
                    }
                    const float cx = xj - bA;
                    const float t = fmaxf(fmaxf(fmaf(-Ui, rA, cx), fmaf(Li, rA, -cx)), 0.f);
                    accA = fmaf(t, t, accA);
                    const float cy = xi - bB;
                    const float u = fmaxf(fmaxf(fmaf(-Uj, rB, cy), fmaf(Lj, rB, -cy)), 0.f);
                    accB = fmaf(u, u, accB);
                };
                auto strided = [&]() __attribute__((always_inline)) {
                    constexpr int S = KS;
                    const int base = 16 * S * kc + kb;
                    // block base pointers, so that the 16 reads per array use immediate offsets
                    // the reads of HB positions go out together (one wait), then their arithmetic
                    constexpr int HB = 8;
                    const float *xaj = L.saj + lj + base, *xai = L.sai + li + base;
                    if constexpr (PW) {
                        const uint32_t *xwi = L.swi + li + base, *xwj = L.swj + lj + base;
#pragma unroll
                        for (int h = 0; h < 16; h += HB) {
                            float va[HB], wa[HB];
                            uint32_t vw[HB], ww[HB];
#pragma unroll
                            for (int t = 0; t < HB; ++t) {
                                va[t] = xaj[S * (h + t)];
                                wa[t] = xai[S * (h + t)];
                                vw[t] = xwi[S * (h + t)];
                                ww[t] = xwj[S * (h + t)];
                            }
                            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                            for (int t = 0; t < HB; ++t) {
                                float Ui, Li, Uj, Lj;
                                L.env(wa[t], vw[t], Ui, Li);
                                L.env(va[t], ww[t], Uj, Lj);
                                if constexpr (FZ) {
                                    fz_term(va[t], wa[t], Ui, Li, Uj, Lj);
                                    continue;
                                }
                                const float cx = va[t] - bA;
                                const float ta = fmaxf(fmaxf(fmaf(-Ui, rA, cx), fmaf(Li, rA, -cx)), 0.f);
                                accA = fmaf(ta, ta, accA);
                                const float cy = wa[t] - bB;
                                const float tb = fmaxf(fmaxf(fmaf(-Uj, rB, cy), fmaf(Lj, rB, -cy)), 0.f);
                                accB = fmaf(tb, tb, accB);
                            }
                        }
                    } else {
                        const float *xui = L.sui + li + base, *xli = L.sli + li + base;
                        const float *xuj = L.suj + lj + base, *xlj = L.slj + lj + base;
#pragma unroll
                        for (int h = 0; h < 16; h += HB) {
                            float va[HB], vu[HB], vl[HB], wa[HB], wu[HB], wl[HB];
#pragma unroll
                            for (int t = 0; t < HB; ++t) {
                                va[t] = xaj[S * (h + t)];
                                vu[t] = xui[S * (h + t)];
                                vl[t] = xli[S * (h + t)];
                                wa[t] = xai[S * (h + t)];
                                wu[t] = xuj[S * (h + t)];
                                wl[t] = xlj[S * (h + t)];
                            }
                            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                            for (int t = 0; t < HB; ++t) {
                                if constexpr (FZ) {
                                    fz_term(va[t], wa[t], vu[t], vl[t], wu[t], wl[t]);
                                    continue;
                                }
                                const float cx = va[t] - bA;
                                const float ta = fmaxf(fmaxf(fmaf(-vu[t], rA, cx), fmaf(vl[t], rA, -cx)), 0.f);
                                accA = fmaf(ta, ta, accA);
                                const float cy = wa[t] - bB;
                                const float tb = fmaxf(fmaxf(fmaf(-wu[t], rB, cy), fmaf(wl[t], rB, -cy)), 0.f);
                                accB = fmaf(tb, tb, accB);
                            }
                        }
                    }
                    if (++kc == nsup) {
                        kc = 0;
                        ++kb;
                    }
                };
                if (k < nstr) {
                    // fixed blocks: 96 independent LDS reads, issued in two groups of 48
                    strided();
                    k += 16;
                    c_kelem += 16ull;
                } else if (k + 16 <= m) {
#pragma unroll
                    for (int kk = 0; kk < 16; ++kk) term(k + kk);
                    k += 16;
                    c_kelem += 16ull;
                } else {
                    c_kelem += (unsigned long long)(m - k);
                    for (; k < m; ++k) term(k);
                }
            }
            const bool dead = act && !((accA <= thrA) && (accB <= thrB));
            const bool done = act && !dead && k >= m;
            if (dead) act = false;
            const unsigned long long dk = __builtin_amdgcn_ballot_w64(done);
            if (dk) {
                if (done) {
                    const int slot = q3n + (int)mbcnt64(dk);
                    q3[slot] = pack_pair(li, lj);
                    if constexpr (FZ) {
                        q3ta[slot] = accA;
                        q3tb[slot] = accB;
                    }
                }
                q3n += __popcll(dk);
                count(C_KEOGH, __popcll(dk));
                if (done) act = false;
                if (q3n > Q3C - 64) break;
            }
        }
        if (q3n > 0) {
            __syncthreads();
            process_q3();
            T2 = th.T2;
            sqmT2 = sqrtf(m * T2);
            sT2 = sqrtf(T2);
        }
        }
        q2n = 0;
    };

    // ---- stage 2b: the segment-mean bound over queued LB_Kim survivors (at most 64 of them),
    // one pair per lane, from the series' table in global memory; survivors go on to LB_Keogh.
    // The entries left (fewer than 64) move to the front of the queue.
    auto process_qp = [&]() __attribute__((always_inline)) {
        const JoinArgs &A = *Ap;
        __syncthreads();  // qp entries were written by other lanes
        const int cnt = qpn < 64 ? qpn : 64;
        const int rest = qpn - cnt;
        bool alive = false;
        uint32_t pr = 0u;
        if (lane < cnt) {
            pr = qp[lane];
            alive = paa_pass(A.paa, cb + pair_row(pr), cb + d0 + pair_col(pr), A.paa_seg, (float)A.paa_len, th.T2);
        }
        const uint32_t tail = lane < rest ? qp[64 + lane] : 0u;
        __syncthreads();  // every lane has read its entry before the queue moves down
        if (lane < rest) qp[lane] = tail;
        qpn = rest;
        const unsigned long long msk = __builtin_amdgcn_ballot_w64(alive);
        if (msk) {
            if (alive) q2[q2n + (int)mbcnt64(msk)] = pr;
            q2n += __popcll(msk);
            count(C_PAA, __popcll(msk));
        }
    };

    // ---- drain: the queue entries (sorted by row) are taken in parts of RC rows: stage the
    // part in LDS, expand its entries into pairs, LB_Kim (stage 2) on 64-lane batches, feeding
    // LB_Keogh and DTW; everything is flushed before the next part is staged.
    auto drain = [&]() __attribute__((always_inline)) {
        const JoinArgs &A = *Ap;
        const CascadeLDS<PW> L(smem, m, GS);
        const int64_t n = A.n, nw = A.nw;
        const int CI = RC + m, CJ = RC + DG + m;
        __syncthreads();  // queue entries were stored by other lanes; column tables retire
        int h0 = 0;
        while (h0 < q1n) {
        // entries [h0, h1) lie in the RC-row part `part` (rows from the item start), which
        // starts at row cb
        const int part = (int)(q1[h0] >> 10) / RC;
        int h1 = h0;
        for (;;) {
            const int idx = h1 + lane;
            const bool in = idx < q1n && (int)(q1[idx] >> 10) / RC == part;
            const int c = __popcll(__builtin_amdgcn_ballot_w64(in));  // a prefix of the 64
            h1 += c;
            if (c < 64) break;
        }
        cb = i_beg + (int64_t)RC * part;
        // Windows of the part's rows (CI samples from cb) and columns (CJ samples from
        // cb + d0), then mean and 1/sigma of its RC rows and RC + DG columns.  Loads go out in
        // batches of 8 per lane (indices clamped into the series, entries past the end set
        // after the load) so that their latencies overlap.
        {
            const global_ptr<float> ac = (global_ptr<float>)A.ac, uc = (global_ptr<float>)A.uc,
                                    lc = (global_ptr<float>)A.lc, muc = (global_ptr<float>)A.mu_c;
            const global_ptr<uint32_t> wc = (global_ptr<uint32_t>)A.wc;
            const global_ptr<float> wf = (global_ptr<float>)A.win;  // WinRow as 8 floats
            constexpr int WISIG = 2;                                // float index of isig
            static_assert(sizeof(WinRow) == 32 && offsetof(WinRow, isig) == 4 * WISIG, "WinRow layout");
            auto stage2 = [&](float *dap, uint32_t *dwp, int64_t base, int cnt) __attribute__((always_inline)) {
                for (int x0 = 0; x0 < cnt; x0 += 512) {
                    float v0[8];
                    uint32_t v1[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int64_t p = base + x0 + 64 * u + lane;
                        const int64_t pc = p < n - 1 ? p : n - 1;
                        v0[u] = ac[pc];
                        v1[u] = wc[pc];
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int x = x0 + 64 * u + lane;
                        if (x < cnt) {
                            const bool ok = base + x < n;
                            dap[x] = ok ? v0[u] : 0.f;
                            dwp[x] = ok ? v1[u] : 0u;
                        }
                    }
                }
            };
            auto stage3 = [&](float *d0p, float *d1p, float *d2p, int64_t base, int cnt) __attribute__((always_inline)) {
                for (int x0 = 0; x0 < cnt; x0 += 512) {
                    float v0[8], v1[8], v2[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int64_t p = base + x0 + 64 * u + lane;
                        const int64_t pc = p < n - 1 ? p : n - 1;
                        v0[u] = ac[pc];
                        v1[u] = uc[pc];
                        v2[u] = lc[pc];
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int x = x0 + 64 * u + lane;
                        if (x < cnt) {
                            const bool ok = base + x < n;
                            d0p[x] = ok ? v0[u] : 0.f;
                            d1p[x] = ok ? v1[u] : 0.f;
                            d2p[x] = ok ? v2[u] : 0.f;
                        }
                    }
                }
            };
            auto stage_stats = [&](float *mup, float *isp, int64_t base, int cnt) __attribute__((always_inline)) {
                for (int x0 = 0; x0 < cnt; x0 += 512) {
                    float v0[8], v1[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int64_t r = base + x0 + 64 * u + lane;
                        const int64_t rc = r < nw - 1 ? r : nw - 1;
                        v0[u] = muc[rc];
                        v1[u] = wf[8 * rc + WISIG];
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int x = x0 + 64 * u + lane;
                        if (x < cnt) {
                            const bool ok = base + x < nw;
                            mup[x] = ok ? v0[u] : 0.f;
                            isp[x] = ok ? v1[u] : 1.f;
                        }
                    }
                }
            };
            if constexpr (PW) {
                stage2(L.sai, L.swi, cb, CI);
                stage2(L.saj, L.swj, cb + d0, CJ);
            } else {
                stage3(L.sai, L.sui, L.sli, cb, CI);
                stage3(L.saj, L.suj, L.slj, cb + d0, CJ);
            }
            stage_stats(L.rmu, L.ris, cb, RC);
            if constexpr (GS) {
                // the part's first column window, cb + d0, lies in the series whenever one of
                // its pairs does; the clamps keep every read inside the series regardless
                const int64_t cg = cb + d0 < nw - 1 ? cb + d0 : nw - 1;
                gcmu = (global_ptr<float>)(A.mu_c + cg);
                gcis = (global_ptr<float>)(A.isg + cg);
                const int64_t lim = nw - 1 - cg;
                gclim = lim < 0 ? 0 : (lim > RC + DG - 1 ? RC + DG - 1 : (int)lim);
            } else {
                stage_stats(L.cmu, L.cis, cb + d0, RC + DG);
            }
        }
        __syncthreads();
        const int row0 = (int)(cb - i_beg);  // part start relative to the item start
        // the next batch's entries and best-so-far word are loaded one batch ahead, so that
        // their latency overlaps the current batch (an older word is only larger: still valid)
        uint32_t e_next = h0 + lane < h1 ? q1[h0 + lane] : 15u;
        unsigned long long key_next = load_bsf(A);
        // Batches of 64 entries; one extra pass after the last batch flushes the LB_Keogh queue,
        // so that process_q2 (and the DTW inside it) is inlined at one place.
        for (int h = h0;; h += 64) {
            const bool last = h >= h1;
            // expand up to 64 entries (row | lane | complemented 4-bit mask; 15 = no entry) into pairs
            const uint32_t e = last ? 15u : e_next;
            const unsigned long long key = key_next;
            e_next = h + 64 + lane < h1 ? q1[h + 64 + lane] : 15u;
            if (!last) key_next = load_bsf(A);
            const uint32_t mask = ~e & 15u;
            const int li = (int)(e >> 10) - row0;
            const int lb = li + KD * (int)((e >> 4) & 63u);
            int total = 0;
            if constexpr (FZ) {
                // pairs listed entry-major: an entry's pairs (consecutive columns of one row)
                // follow those of the lanes below it, so that neighbouring list slots read
                // neighbouring windows in the stages behind (fewer LDS bank conflicts).  The
                // offsets are the prefix sums of the per-lane counts (0..4), bit by bit.
                const uint32_t cnt = (uint32_t)__popc(mask);
                const unsigned long long c0 = __builtin_amdgcn_ballot_w64(cnt & 1u);
                const unsigned long long c1 = __builtin_amdgcn_ballot_w64(cnt & 2u);
                const unsigned long long c2 = __builtin_amdgcn_ballot_w64(cnt & 4u);
                int o = (int)mbcnt64(c0) + 2 * (int)mbcnt64(c1) + 4 * (int)mbcnt64(c2);
                total = __popcll(c0) + 2 * __popcll(c1) + 4 * __popcll(c2);
#pragma unroll
                for (int q = 0; q < KD; ++q) {
                    if ((mask >> q) & 1u) plist[o++] = pack_pair(li, lb + q);
                }
            } else {
                // pairs listed diagonal-bit-major: the pairs of bit q follow those of bits < q,
                // in lane order (one ballot per bit instead of a cross-lane prefix sum)
#pragma unroll
                for (int q = 0; q < KD; ++q) {
                    const bool has = (mask >> q) & 1u;
                    const unsigned long long bq = __builtin_amdgcn_ballot_w64(has);
                    if (has) plist[total + (int)mbcnt64(bq)] = pack_pair(li, lb + q);
                    total += __popcll(bq);
                }
            }
            count(C_BALL, (unsigned long long)total);
            __syncthreads();
            set_th(key);
            for (int p = 0;; p += 64) {
                const bool act = p + lane < total;
                int pi = 0, pj = 0;
                bool alive = false;
                if (act) {
                    const uint32_t pr = plist[p + lane];
                    pi = pair_row(pr);
                    pj = pair_col(pr);
                    const float *ai = L.sai + pi, *aj = L.saj + pj;
                    // the 16 LDS reads of the pair go out together (one wait), then the arithmetic
                    const float mi = L.rmu[pi], ii = L.ris[pi], mj = col_mu(L, pj), ij = col_is(L, pj);
                    const float a0 = ai[0], a1 = ai[1], a2 = ai[2], b0 = aj[0], b1 = aj[1], b2 = aj[2];
                    const float ae0 = ai[m - 1], ae1 = ai[m - 2], ae2 = ai[m - 3];
                    const float be0 = aj[m - 1], be1 = aj[m - 2], be2 = aj[m - 3];
                    __builtin_amdgcn_sched_barrier(0);
                    // Any warping path from (0,0) first leaves {0,1}^2 at (2,0),(2,1),(2,2),(1,2) or
                    // (0,2), entered from an adjacent cell of {0,1}^2.  Same at the far corner.
                    auto corner = [](float c00, float c10, float c01, float c11, float c20, float c21, float c22,
                                     float c12, float c02) __attribute__((always_inline)) {
                        float b = c10 + c20;
                        b = fminf(b, fminf(c10, c11) + c21);
                        b = fminf(b, c11 + c22);
                        b = fminf(b, fminf(c01, c11) + c12);
                        b = fminf(b, c01 + c02);
                        return c00 + b;
                    };
                    // x: window i (columns), y: window j (rows)
                    float kim;
                    if constexpr (FZ) {
                        // the DTW's own fused difference fma(y, -1/sigma_j, fma(x, 1/sigma_i, c))
                        // (within the 2 eps_z of two rounded z-values; the row term formed once
                        // per sample of window i)
                        const float cz = pair_bias(mi, ii, mj, ij), nij = -ij;
                        const float x0 = fmaf(a0, ii, cz), x1 = fmaf(a1, ii, cz), x2 = fmaf(a2, ii, cz);
                        const float xe0 = fmaf(ae0, ii, cz), xe1 = fmaf(ae1, ii, cz), xe2 = fmaf(ae2, ii, cz);
                        auto d2 = [&](float y, float xz) __attribute__((always_inline)) { return sq(fmaf(y, nij, xz)); };
                        kim = corner(d2(b0, x0), d2(b1, x0), d2(b0, x1), d2(b1, x1), d2(b2, x0), d2(b2, x1), d2(b2, x2),
                                     d2(b1, x2), d2(b0, x2)) +
                              corner(d2(be0, xe0), d2(be1, xe0), d2(be0, xe1), d2(be1, xe1), d2(be2, xe0), d2(be2, xe1),
                                     d2(be2, xe2), d2(be1, xe2), d2(be0, xe2));
                    } else {
                        // z-values as one FMA each, x / sigma - mean / sigma (within eps_z)
                        const float nmi = -mi * ii, nmj = -mj * ij;
                        const float x0 = fmaf(a0, ii, nmi), x1 = fmaf(a1, ii, nmi), x2 = fmaf(a2, ii, nmi);
                        const float y0 = fmaf(b0, ij, nmj), y1 = fmaf(b1, ij, nmj), y2 = fmaf(b2, ij, nmj);
                        const float xe0 = fmaf(ae0, ii, nmi), xe1 = fmaf(ae1, ii, nmi), xe2 = fmaf(ae2, ii, nmi);
                        const float ye0 = fmaf(be0, ij, nmj), ye1 = fmaf(be1, ij, nmj), ye2 = fmaf(be2, ij, nmj);
                        const float c00 = sq(y0 - x0), c10 = sq(y1 - x0), c01 = sq(y0 - x1), c11 = sq(y1 - x1);
                        const float c20 = sq(y2 - x0), c21 = sq(y2 - x1), c22 = sq(y2 - x2), c12 = sq(y1 - x2),
                                    c02 = sq(y0 - x2);
                        float bl = c10 + c20;
                        bl = fminf(bl, fminf(c10, c11) + c21);
                        bl = fminf(bl, c11 + c22);
                        bl = fminf(bl, fminf(c01, c11) + c12);
                        bl = fminf(bl, c01 + c02);
                        const float e00 = sq(ye0 - xe0), e10 = sq(ye1 - xe0), e01 = sq(ye0 - xe1), e11 = sq(ye1 - xe1);
                        const float e20 = sq(ye2 - xe0), e21 = sq(ye2 - xe1), e22 = sq(ye2 - xe2), e12 = sq(ye1 - xe2),
                                    e02 = sq(ye0 - xe2);
                        float br = e10 + e20;
                        br = fminf(br, fminf(e10, e11) + e21);
                        br = fminf(br, e11 + e22);
                        br = fminf(br, fminf(e01, e11) + e12);
                        br = fminf(br, e01 + e02);
                        kim = (c00 + bl) + (e00 + br);
                    }
                    alive = kim <= th.kim;
                }
                const unsigned long long msk = __builtin_amdgcn_ballot_w64(alive);
                const bool paa = A.paa_seg > 0;
                if (msk) {
                    if (paa) {
                        if (alive) qp[qpn + (int)mbcnt64(msk)] = pack_pair(pi, pj);
                        qpn += __popcll(msk);
                    } else {
                        if (alive) q2[q2n + (int)mbcnt64(msk)] = pack_pair(pi, pj);
                        q2n += __popcll(msk);
                    }
                    count(C_KIM, __popcll(msk));
                }
                // queued pairs refer to this part's staging: all of them run before the next part
                if (paa && (qpn >= 64 || (last && qpn > 0))) process_qp();
                if (q2n >= Q2C - 64 || (last && q2n > 0)) process_q2();
                if (p + 64 >= total) break;
            }
            __syncthreads();  // plist is rewritten by the next batch
            if (last) break;
        }
        h0 = h1;
        }
        q1n = 0;
        __syncthreads();
    };

    for (;;) {
        // 64-bit claim counter: the item count of a long series exceeds 2^31
        unsigned long long item_l = 0;
        if (lane == 0) item_l = atomicAdd(work_counter, 1ull);
        const int64_t claim = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(item_l >> 32)) << 32) |
                                        (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)item_l));
        const int64_t total = item_base[nseries];
        if (claim >= total) break;
        // claim order: claim k takes item k * mul mod total (mul coprime with total, a bijection),
        // so that the first claims are spread over the whole pair space and the best-so-far falls
        // early wherever the best pairs lie; mul = 0 keeps the index order
        const uint64_t mul = (uint64_t)item_base[nseries + 1];
        const int64_t gitem = mul ? (int64_t)(((uint64_t)claim * mul) % (uint64_t)total) : claim;
        // series of this item (usually the one of the previous item)
        int s = cur;
        if (s < 0 || gitem < item_base[s] || gitem >= item_base[s + 1]) {
            int lo = 0, hi = nseries - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (item_base[mid] <= gitem)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            s = lo;
        }
        if (s != cur) {
            flush_counters();
            cur = s;
            Ap = series + s;
            m = Ap->m;
            thk = ~0ull;
        }
        const JoinArgs &A = *Ap;
        if (!A.valid) continue;
        // Sweep test for this item: the one-sided ball bound (X^A only, half the work per pair)
        // when the previous item of this series left almost no survivors; the two-sided one
        // otherwise.  Both are valid bounds; the cascade's two-sided LB_Keogh follows either.
        if (A.sweep_mode != 0) {  // fixed by the caller
            two_sided = A.sweep_mode == 1;
            prev_series = s;
        } else if (s != prev_series) {
            two_sided = true;
            prev_series = s;
        } else if (prev_rows > 0) {
            const float rate = (float)item_entries / (64.0f * (float)prev_rows);  // entries per lane-row
            two_sided = two_sided ? !(rate < 1e-3f) : (rate > 4e-3f);
        }
        item_entries = 0;
        const int64_t nw = A.nw, n = A.n;
        const int64_t item = gitem - item_base[s];
        int g;
        int64_t sidx;
        decode_item(A.item_start, A.groups, item, g, sidx);
        d0 = (int64_t)m + (int64_t)DG * g;
        const int64_t rows_g = nw - d0;  // rows of diagonal d0
        i_beg = sidx * A.rows_per_item;
        const int64_t i_end = (i_beg + A.rows_per_item < rows_g) ? i_beg + A.rows_per_item : rows_g;
        prev_rows = i_end - i_beg;
        const int64_t dq0 = d0 + KD * lane;
        // last valid row per owned diagonal: i <= nw - 1 - d
        int64_t lim[KD];
#pragma unroll
        for (int q = 0; q < KD; ++q) lim[q] = nw - 1 - (dq0 + q);

        // ---- exact FP64 anchor of X^A, X^B at row i_beg: for the lane's four consecutive
        // columns j = i_beg + dq0 + q, X^A_q = sum_k (v_{i+k} - muv_i) a_{j+k} and
        // X^B_q = sum_k (a_{i+k} - mu_i) v_{j+k}.  Two passes (X^A, then X^B) over windows staged
        // in LDS (the union area is free at this point) with coalesced loads: the row window
        // (m values) and the column span (DG + m + 3 values from i_beg + d0).  Positions past
        // the series end read as 0; they only reach columns past the triangle, never used.
        float XA[KD], XB[KD];
#pragma unroll
        for (int q = 0; q < KD; ++q) XB[q] = 0.f;
        {
            double *lrow = reinterpret_cast<double *>(smem);
            double *lcol = lrow + m;
            const int span = DG + m + KD;
            const int64_t jb0 = i_beg + d0;
#pragma unroll 1
            for (int pass = 0; pass < (two_sided ? 2 : 1); ++pass) {
                const global_ptr<double> prow = in_vgprs(pass == 0 ? A.v : A.a);
                const global_ptr<double> pcol = in_vgprs(pass == 0 ? A.a : A.v);
                const double c0v = pass == 0 ? A.muv[i_beg] : A.mu[i_beg];
                __syncthreads();
                for (int x0 = 0; x0 < m; x0 += 256) {
                    double t4[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int x = x0 + 64 * u + lane;
                        t4[u] = x < m ? prow[i_beg + x] - c0v : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int x = x0 + 64 * u + lane;
                        if (x < m) lrow[x] = t4[u];
                    }
                }
                for (int x0 = 0; x0 < span; x0 += 256) {
                    double t4[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int64_t p = jb0 + x0 + 64 * u + lane;
                        t4[u] = (x0 + 64 * u + lane < span && p < n) ? pcol[p] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int x = x0 + 64 * u + lane;
                        if (x < span) lcol[x] = t4[u];
                    }
                }
                __syncthreads();
                double acc[KD];
#pragma unroll
                for (int q = 0; q < KD; ++q) acc[q] = 0.0;
                const double *cl = lcol + KD * lane;
                double win[KD];
#pragma unroll
                for (int q = 0; q < KD - 1; ++q) win[q + 1] = cl[q];
#pragma unroll 8
                for (int k = 0; k < m; ++k) {
#pragma unroll
                    for (int q = 0; q < KD - 1; ++q) win[q] = win[q + 1];
                    win[KD - 1] = cl[k + KD - 1];
                    const double ck = lrow[k];
#pragma unroll
                    for (int q = 0; q < KD; ++q) acc[q] = fma(ck, win[q], acc[q]);
                }
#pragma unroll
                for (int q = 0; q < KD; ++q) {
                    if (pass == 0)
                        XA[q] = (float)acc[q];
                    else
                        XB[q] = (float)acc[q];
                }
            }
            __syncthreads();
        }
        // pairs of the item inside the triangle: row i has min(DG, nw - d0 - i) >= 1 columns
        {
            const int64_t F = nw - d0;
            const int64_t full_end = (F - DG + 1 < i_end) ? F - DG + 1 : i_end;  // rows with DG columns
            int64_t cnt = full_end > i_beg ? (full_end - i_beg) * DG : 0;
            const int64_t a0 = (i_beg > F - DG + 1) ? i_beg : F - DG + 1, b0 = i_end - 1;
            if (a0 <= b0) cnt += (b0 - a0 + 1) * (2 * F - a0 - b0) / 2;
            count(C_SWEPT, (unsigned long long)cnt);
            if (!two_sided) count(C_SWEPT_ONE, (unsigned long long)cnt);
        }
        // ---- per-item rounding margins of the FP32 sweep, in the units of X / sigma_j
        // (resp. X / sigma_i): each step rounds two FMAs on values bounded by
        // |X| <= m sigma_j sv_i, and the FP32 inputs carry 2^-24 relative error.
        const int64_t jlo = i_beg + d0, jhi = (i_end - 1 + d0 + DG - 1 < nw - 1) ? i_end - 1 + d0 + DG - 1 : nw - 1;
        const float smaxJ = range_max(A, 0, jlo, jhi), svmaxJ = range_max(A, 1, jlo, jhi),
                    ismaxJ = range_max(A, 2, jlo, jhi);
        const float smaxI = range_max(A, 0, i_beg, i_end - 1), svmaxI = range_max(A, 1, i_beg, i_end - 1),
                    ismaxI = range_max(A, 2, i_beg, i_end - 1);
        const float steps = (float)(i_end - i_beg + 8);
        const float EA = 0x1p-22f * steps * (float)m * smaxJ * svmaxI * ismaxJ;
        const float EB = 0x1p-22f * steps * (float)m * smaxI * svmaxJ * ismaxI;

        // LDS tables of the sweep.  Two-sided: NJARR column tables of JW2 entries and NJARR row
        // tables of RC2 entries.  One-sided: 3 + 3 tables of chunks RC1 rows high (columns dfa dga
        // 1/sigma, rows dfv dgv thetaA).  Either chunk is drained in RC-row parts.
        const int rc = two_sided ? RC2 : RC1;
        float *const rarr1 = jarr + 3 * JW1;
        for (c0 = i_beg; c0 < i_end; c0 += rc) {
            const int nrows = (int)(((c0 + rc < i_end) ? c0 + rc : i_end) - c0);
            // rows of the chunk whose pair with column q is inside the triangle: r <= limr[q]
            int limr[KD];
#pragma unroll
            for (int q = 0; q < KD; ++q) {
                const int64_t l = lim[q] - c0;
                limr[q] = l < -1 ? -1 : (l > rc ? rc : (int)l);
            }
            // the whole chunk is inside the triangle for every lane
            const bool chunk_full = c0 + nrows - 1 + d0 + DG - 1 <= nw - 1;
            const uint32_t code0 = (uint32_t)(c0 - i_beg) << 10;
            int rl = 0;  // next local row (multiple of 8)
            while (rl < nrows) {
                // Stage the chunk's row data (rows c0 + x) and column data (columns c0 + d0 + x)
                // into LDS: two-sided, dfa dga dfv dgv 1/sigma and thetaA / thetaB; one-sided,
                // columns dfa dga 1/sigma (the first 16 bytes of a window record) and rows dfv
                // dgv thetaA.  Every window load is issued before the first use (indices clamped
                // into the series, entries past the end zeroed after the load), so that the
                // latencies of the chunk's loads overlap instead of adding up.
                const global_ptr<vf4> win = (global_ptr<vf4>)A.win;
                static_assert(sizeof(WinRow) == 32 && offsetof(WinRow, isig) == 8 && offsetof(WinRow, dfv) == 12,
                              "WinRow is read as two 16-byte vectors");
                // A staged threshold is never +0 (+0 is stored as -0, the same number): the test
                // value fma(X, 1/sigma, -theta) is then never -0, so its sign bit alone says
                // whether the pair is below the threshold (the queue mask is built from sign bits).
                auto theta = [&](const vf4 &lo, const vf4 &hi, float T, float E) __attribute__((always_inline)) {
                    // (mN, R, sig) = hi.y, hi.z, hi.w
                    const float t = T + hi.z;
                    return no_pos_zero((0.5f * (hi.y - t * t) - 0x1p-20f * (hi.y + t * t)) * hi.w - E);
                };
                if (two_sided) {
                    constexpr int NRL = RC2 / 64, NCL = (JW2 + 63) / 64;
                    vf4 rlo[NRL], rhi[NRL], clo[NCL], chi[NCL];
#pragma unroll
                    for (int k = 0; k < NRL; ++k) {
                        const int64_t i = c0 + lane + 64 * k, ic = i < nw - 1 ? i : nw - 1;
                        rlo[k] = win[2 * ic];
                        rhi[k] = win[2 * ic + 1];
                    }
#pragma unroll
                    for (int k = 0; k < NCL; ++k) {
                        const int64_t j = c0 + d0 + lane + 64 * k, jc = j < nw - 1 ? j : nw - 1;
                        clo[k] = win[2 * jc];
                        chi[k] = win[2 * jc + 1];
                    }
                    th = load_threshold(A);
#pragma unroll
                    for (int k = 0; k < NRL; ++k) {
                        const int x = lane + 64 * k;
                        const bool ok = x < nrows;
                        rarr[0 * RC2 + x] = ok ? rlo[k].x : 0.f;        // dfa
                        rarr[1 * RC2 + x] = ok ? rlo[k].y : 0.f;        // dga
                        rarr[2 * RC2 + x] = ok ? rlo[k].w : 0.f;        // dfv
                        rarr[3 * RC2 + x] = ok ? rhi[k].x : 0.f;        // dgv
                        rarr[4 * RC2 + x] = ok ? rlo[k].z : 0.f;        // 1/sigma
                        rarr[5 * RC2 + x] = ok ? theta(rlo[k], rhi[k], th.T, EA) : -0.f;
                    }
#pragma unroll
                    for (int k = 0; k < NCL; ++k) {
                        const int x = lane + 64 * k;
                        if (x < JW2) {
                            const bool ok = c0 + d0 + x <= nw - 1;
                            jarr[0 * JW2 + x] = ok ? clo[k].x : 0.f;
                            jarr[1 * JW2 + x] = ok ? clo[k].y : 0.f;
                            jarr[2 * JW2 + x] = ok ? clo[k].w : 0.f;
                            jarr[3 * JW2 + x] = ok ? chi[k].x : 0.f;
                            jarr[4 * JW2 + x] = ok ? clo[k].z : 0.f;
                            jarr[5 * JW2 + x] = ok ? theta(clo[k], chi[k], th.T, EB) : -0.f;
                        }
                    }
                } else {
                    constexpr int NRL = RC1 / 64, NCL = (JW1 + 63) / 64;
                    vf4 rlo[NRL], rhi[NRL], clo[NCL];
#pragma unroll
                    for (int k = 0; k < NRL; ++k) {
                        const int64_t i = c0 + lane + 64 * k, ic = i < nw - 1 ? i : nw - 1;
                        rlo[k] = win[2 * ic];
                        rhi[k] = win[2 * ic + 1];
                    }
#pragma unroll
                    for (int k = 0; k < NCL; ++k) {
                        const int64_t j = c0 + d0 + lane + 64 * k, jc = j < nw - 1 ? j : nw - 1;
                        clo[k] = win[2 * jc];
                    }
                    th = load_threshold(A);
#pragma unroll
                    for (int k = 0; k < NRL; ++k) {
                        const int x = lane + 64 * k;
                        const bool ok = x < nrows;
                        rarr1[0 * RC1 + x] = ok ? rlo[k].w : 0.f;  // dfv
                        rarr1[1 * RC1 + x] = ok ? rhi[k].x : 0.f;  // dgv
                        rarr1[2 * RC1 + x] = ok ? theta(rlo[k], rhi[k], th.T, EA) : -0.f;
                    }
#pragma unroll
                    for (int k = 0; k < NCL; ++k) {
                        const int x = lane + 64 * k;
                        if (x < JW1) {
                            const bool ok = c0 + d0 + x <= nw - 1;
                            jarr[0 * JW1 + x] = ok ? clo[k].x : 0.f;  // dfa
                            jarr[1 * JW1 + x] = ok ? clo[k].y : 0.f;  // dga
                            jarr[2 * JW1 + x] = ok ? clo[k].z : 0.f;  // 1/sigma
                        }
                    }
                }
                __syncthreads();
                // Column slots of this lane: slot q holds the staged column x = r + 4 lane + q
                // for the current local row r.  Moving to row r + 1 shifts every slot by one
                // column: slots 1..3 move down (register renaming in the unrolled group) and
                // slot 3 takes column r + 4 lane + 4.  For the 4 rows of a half starting at r0
                // (a multiple of 4) the entering columns are one aligned float4 per table at
                // x = r0 + 4 (lane + 1), read from LDS (LDS pipe, no VALU shifts).
                // Table r (dfa dga dfv dgv 1/sigma theta) of the mode's layout.
                auto colp = [&](auto two_tag, int r) __attribute__((always_inline)) {
                    return decltype(two_tag)::value ? jarr + r * JW2 : jarr + (r == 4 ? 2 : r) * JW1;
                };
                auto rowp = [&](auto two_tag, int r) __attribute__((always_inline)) {
                    return decltype(two_tag)::value ? rarr + r * RC2 : rarr1 + (r == 5 ? 2 : r - 2) * RC1;
                };
                float cs[NJARR][KD];
                // One 8-row group: rows rl..rl+7 of the chunk, in two halves of 4 rows whose
                // entering columns are read into separate registers, so that the rotating
                // column slots end every group in the registers they started in.  FULL: every
                // pair is inside the triangle.  Test per pair: X^A / sigma_j >= thetaA_i and
                // (TWO) X^B / sigma_i >= thetaB_j as signs of single-rounding FMAs; a lane-wide
                // min/max reduces the four pairs to one compare per row, and a queue entry
                // (row | lane | sign bits of the lane's four test values) is only formed on rows
                // with a survivor.  Without TWO only X^A and its three tables are carried.
                auto group = [&](auto full_tag, auto two_tag) __attribute__((always_inline)) {
                    constexpr bool FULL = decltype(full_tag)::value;
                    constexpr bool TWO = decltype(two_tag)::value;
                    // tables used: columns dfa dga (1/sigma) [dfv dgv thetaB]; rows dfv dgv
                    // thetaA [dfa dga 1/sigma]
                    auto colused = [](int r) { return TWO || r == 0 || r == 1 || r == 4; };
                    auto rowused = [](int r) { return TWO || r == 2 || r == 3 || r == 5; };
                    (void)full_tag;
                    // entry of this lane for the group's first row (rows add their offset << 10)
                    const uint32_t ebase = code0 + ((uint32_t)rl << 10) + ((uint32_t)lane << 4);
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int r0 = rl + 4 * h;
                        if (!FULL && r0 >= nrows) break;
                        float rv[NJARR][4];  // row data of the 4 rows (same address on every lane)
                        float nv[NJARR][4];  // columns entering slot 3: x = r0 + 4 (lane + 1) + u
#pragma unroll
                        for (int r = 0; r < NJARR; ++r) {
                            if (rowused(r)) {
                                const float4 v4 = *reinterpret_cast<const float4 *>(rowp(two_tag, r) + r0);
                                rv[r][0] = v4.x;
                                rv[r][1] = v4.y;
                                rv[r][2] = v4.z;
                                rv[r][3] = v4.w;
                            }
                            if (colused(r)) {
                                const float4 n4 =
                                    *reinterpret_cast<const float4 *>(colp(two_tag, r) + r0 + 4 * lane + 4);
                                nv[r][0] = n4.x;
                                nv[r][1] = n4.y;
                                nv[r][2] = n4.z;
                                nv[r][3] = n4.w;
                            }
                        }
                        // queue entries of row r0 + u: one per lane with a surviving diagonal
                        float t[4][KD];
                        auto append = [&](int u) __attribute__((always_inline)) {
                            const float tmax = fmaxf(fmaxf(t[u][0], t[u][1]), fmaxf(t[u][2], t[u][3]));
                            const bool any = tmax >= 0.0f;
                            const unsigned long long mk = __builtin_amdgcn_ballot_w64(any);
                            if (__builtin_expect(mk != 0ull, 0)) {
                                if (any) {
                                    // the entry stores the complement of the survivor mask: bit q
                                    // is the sign bit of t (t is never -0, so a clear sign bit
                                    // means t >= 0; a NaN test value can only add a pair)
                                    const uint32_t neg = (__float_as_uint(t[u][0]) >> 31) |
                                                         ((__float_as_uint(t[u][1]) >> 30) & 2u) |
                                                         ((__float_as_uint(t[u][2]) >> 29) & 4u) |
                                                         ((__float_as_uint(t[u][3]) >> 28) & 8u);
                                    q1[(uint32_t)q1n + mbcnt64(mk)] = ebase + ((uint32_t)(4 * h + u) << 10) + neg;
                                }
                                q1n += __popcll(mk);
                                item_entries += __popcll(mk);
                            }
                        };
                        // one-sided full halves test all 16 pairs with one ballot (survivors are
                        // rare there by construction) and form entries row by row only on a hit
                        constexpr bool BATCH = FULL && !TWO;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            if (!FULL && r0 + u >= nrows) break;
#pragma unroll
                            for (int q = 0; q < KD; ++q) {
                                const float sA = fmaf(XA[q], cs[4][q], -rv[5][u]);
                                if (TWO) {
                                    const float sB = fmaf(XB[q], rv[4][u], -cs[5][q]);
                                    t[u][q] = fminf(sA, sB);
                                } else {
                                    t[u][q] = sA;
                                }
                                if (!FULL && r0 + u > limr[q]) t[u][q] = -1.0f;
                            }
                            if (!BATCH) append(u);
#pragma unroll
                            for (int q = 0; q < KD; ++q) {
                                XA[q] = fmaf(cs[0][q], rv[3][u], XA[q]);
                                XA[q] = fmaf(rv[2][u], cs[1][q], XA[q]);
                                if (TWO) {
                                    XB[q] = fmaf(rv[0][u], cs[3][q], XB[q]);
                                    XB[q] = fmaf(cs[2][q], rv[1][u], XB[q]);
                                }
                            }
#pragma unroll
                            for (int r = 0; r < NJARR; ++r) {
                                if (!colused(r)) continue;
                                cs[r][0] = cs[r][1];
                                cs[r][1] = cs[r][2];
                                cs[r][2] = cs[r][3];
                                cs[r][3] = nv[r][u];
                            }
                        }
                        if (BATCH) {
                            float hmax = fmaxf(fmaxf(t[0][0], t[0][1]), t[0][2]);
#pragma unroll
                            for (int k = 3; k + 1 < 4 * KD; k += 2)  // pairs (3,4) .. (13,14)
                                hmax = fmaxf(fmaxf(hmax, t[k / KD][k % KD]), t[(k + 1) / KD][(k + 1) % KD]);
                            hmax = fmaxf(hmax, t[3][KD - 1]);
                            if (__builtin_expect(__builtin_amdgcn_ballot_w64(hmax >= 0.0f) != 0ull, 0)) {
#pragma unroll
                                for (int u = 0; u < 4; ++u) append(u);
                            }
                        }
                    }
                };
                using T_ = std::integral_constant<bool, true>;
                using F_ = std::integral_constant<bool, false>;
                auto sweep = [&](auto two_tag) __attribute__((always_inline)) {
                    constexpr bool TWO = decltype(two_tag)::value;
#pragma unroll
                    for (int r = 0; r < NJARR; ++r) {
                        if (!(TWO || r == 0 || r == 1 || r == 4)) continue;
                        const float4 v4 = *reinterpret_cast<const float4 *>(colp(two_tag, r) + rl + 4 * lane);
                        cs[r][0] = v4.x;
                        cs[r][1] = v4.y;
                        cs[r][2] = v4.z;
                        cs[r][3] = v4.w;
                    }
                    // whole 8-row groups of a chunk inside the triangle take the unchecked form;
                    // a queue close to full ends the sweep (the chunk resumes after the drain)
                    bool stop = false;
                    if (chunk_full) {
                        const int nfull = nrows & ~7;
                        for (; rl < nfull; rl += 8) {
                            group(T_(), two_tag);
                            if (q1n > QLIM) {
                                rl += 8;
                                stop = true;
                                break;
                            }
                        }
                    }
                    if (!stop)
                        for (; rl < nrows; rl += 8) {
                            group(F_(), two_tag);
                            if (q1n > QLIM) {
                                rl += 8;
                                break;
                            }
                        }
                };
                if (two_sided)
                    sweep(T_());
                else
                    sweep(F_());
                // single drain point (end of chunk, or before the queue could overflow)
                if (q1n > 0) drain();
                __syncthreads();
            }
        }
    }
    flush_counters();
}

template <int Q, bool PW, bool SW>
__global__ __launch_bounds__(64) void k_join(const JoinArgs *__restrict__ series, const int64_t *__restrict__ item_base,
                                             int32_t nseries, int32_t uni, unsigned long long *work_counter, uint32_t *queue) {
    join_body<Q, PW, SW>(series, item_base, nseries, uni, work_counter, queue);
}

// The same join held to the register budget of two waves per SIMD (a few spills to scratch):
// used for bands of two offset pairs per lane, whose packed-envelope staging lets two waves
// per SIMD fit the LDS up to m = 1024.
template <int Q, bool GS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_join_occ2(
    const JoinArgs *__restrict__ series, const int64_t *__restrict__ item_base, int32_t nseries, int32_t uni,
    unsigned long long *work_counter, uint32_t *queue) {
    join_body<Q, true, false, GS>(series, item_base, nseries, uni, work_counter, queue);
}

// ---- set-up of every series of a join: the FP32 error model from the preprocessing's value
// range, the validity flag, and the best-so-far / candidate / counter resets (a kernel, so
// that a captured graph holds no memset node).  One thread per series.
__global__ void k_join_setup(JoinArgs *series, int32_t nseries, unsigned long long *work_counter) {
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s == 0) *work_counter = 0ull;
    if (s >= nseries) return;
    JoinArgs &A = series[s];
    const float *g = A.gstat;
    const double m = (double)A.m;
    A.amax = g[0];
    A.eps_z = (float)(0x1p-22 * (((double)g[0] + (double)g[1]) * (double)g[2] + sqrt(m)));
    A.rel_dtw = (float)(0x1p-10 + 4.0 * m * 0x1p-24);
    A.valid = (A.flags[0] == 0 && A.flags[1] == 0) ? 1 : 0;
    *A.bsf = A.init_key;
    *A.n_cands = 0ull;
    for (int c = 0; c < C_NCOUNTERS; ++c) A.counters[c] = 0ull;
}

// ---- seed scan: z-normalised Euclidean distance (an upper bound of banded DTW) over a
// sample of work items; each wave reports its closest pair (PER_LANE: every lane reports the
// closest pair of its own four diagonals, the candidate source of the learned proposal).
// Block n_seed writes the fixed seed pair (0, m).
template <bool PER_LANE>
__device__ __forceinline__ void seed_block(const SeedArgs &S, const int blk, int2 *pairs) {
    const int lane = (int)threadIdx.x;
    const int m = S.m;
    if (blk == S.n_seed) {
        if (lane == 0 && !PER_LANE) pairs[blk] = make_int2(0, m);
        return;
    }
    const int64_t nw = S.nw;
    const int64_t item = ((int64_t)blk * S.n_items) / S.n_seed;
    int g;
    int64_t sidx;
    decode_item(S.item_start, S.groups, item, g, sidx);
    const int64_t d0 = (int64_t)m + (int64_t)DG * g;
    const int64_t rows_g = nw - d0;
    const int64_t i_beg = sidx * (int64_t)S.rows_per_item;
    int64_t i_end = i_beg + S.rows;
    if (i_end > rows_g) i_end = rows_g;
    const int64_t dq0 = d0 + KD * lane;
    // FP64 anchor of the lane's four diagonals at row i_beg: the column values slide through
    // registers (one new load per k, reads past the series end only feed diagonals outside
    // the triangle, which are discarded); loads run 8 positions ahead of their use.
    float C[KD];
    {
        const int64_t n = S.n;
        const global_ptr<double> pa = (global_ptr<double>)S.a;
        const int64_t jb = i_beg + dq0;
        auto col = [&](int64_t p) { return pa[p < n - 1 ? p : n - 1]; };
        double c[KD] = {0.0, 0.0, 0.0, 0.0};
        const double m0 = S.mu[i_beg];
        double win[KD];
#pragma unroll
        for (int q = 0; q < KD - 1; ++q) win[q + 1] = col(jb + q);
#pragma unroll 8
        for (int k = 0; k < m; ++k) {
            const double cc = pa[i_beg + k] - m0;
#pragma unroll
            for (int q = 0; q < KD - 1; ++q) win[q] = win[q + 1];
            win[KD - 1] = col(jb + k + KD - 1);
#pragma unroll
            for (int q = 0; q < KD; ++q) c[q] = fma(cc, win[q], c[q]);
        }
#pragma unroll
        for (int q = 0; q < KD; ++q) C[q] = i_beg + dq0 + q <= nw - 1 ? (float)c[q] : 0.f;
    }
    // rows in batches of 4 whose window records (the 16-byte halves holding dfa, dga, 1/sigma)
    // are all loaded before the first use; indices clamped into the table, out-of-range pairs
    // skipped exactly as in a row-by-row scan
    const global_ptr<vf4> win4 = (global_ptr<vf4>)S.win;
    static_assert(sizeof(WinRow) == 32 && offsetof(WinRow, isig) == 8, "WinRow is read as two 16-byte vectors");
    float best = __builtin_huge_valf();
    int bi = -1, bj = -1;
    for (int64_t i0 = i_beg; i0 < i_end; i0 += 4) {
        vf4 ri[4], rj[4][KD];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t i = i0 + u < nw - 1 ? i0 + u : nw - 1;
            ri[u] = win4[2 * i];
#pragma unroll
            for (int q = 0; q < KD; ++q) {
                const int64_t j = i0 + u + dq0 + q, jc = j < nw - 1 ? j : nw - 1;
                rj[u][q] = win4[2 * jc];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t i = i0 + u;
            if (i >= i_end) break;
#pragma unroll
            for (int q = 0; q < KD; ++q) {
                const int64_t j = i + dq0 + q;
                if (j <= nw - 1) {
                    // (dfa, dga, isig) = .x, .y, .z
                    const float ed2 = 2.0f * m - 2.0f * C[q] * ri[u].z * rj[u][q].z;
                    if (ed2 < best) {
                        best = ed2;
                        bi = (int)i;
                        bj = (int)j;
                    }
                    C[q] = C[q] + rj[u][q].x * ri[u].y + ri[u].x * rj[u][q].y;
                }
            }
        }
    }
    if (PER_LANE) {
        pairs[(size_t)blk * 64 + lane] = make_int2(bi, bj);
        return;
    }
    // wave argmin
    float v = best;
    for (int off = 32; off > 0; off >>= 1) {
        float o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    const unsigned long long msk = __ballot(best == v && bi >= 0);
    const int src = msk ? (int)__ffsll((long long)msk) - 1 : 0;
    const int wi = __shfl(bi, src, 64), wj = __shfl(bj, src, 64);
    if (lane == 0) pairs[blk] = msk ? make_int2(wi, wj) : make_int2(-1, -1);
}

template <bool PER_LANE>
__global__ __launch_bounds__(64) void k_seed(SeedArgs S, int2 *pairs) {
    seed_block<PER_LANE>(S, (int)blockIdx.x, pairs);
}

// Series t of a batch owns blocks base[t] .. base[t+1] - 1 (first t with base[t+1] > b).
__device__ inline int batch_owner(const int32_t *base, int32_t nj, int32_t b) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (base[mid + 1] > b)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// The seed scans of every series of a batch in one launch.
__global__ __launch_bounds__(64) void k_seed_batch(const SeedJob *__restrict__ jobs, const int32_t *__restrict__ base,
                                                   int32_t nj) {
    const int b = (int)blockIdx.x;
    const int t = batch_owner(base, nj, b);
    seed_block<false>(jobs[t].S, b - base[t], jobs[t].pairs);
}

// ---- DTW of a list of candidate pairs (seeds), one pair per wave; pushes results like the
// join does.
template <int QP>
__device__ __forceinline__ void seed_dtw_block(const JoinArgs &A, const int2 *pairs, int npairs, int p, float *zs) {
    const int lane = (int)threadIdx.x;
    const int m = A.m;
    if (p >= npairs || !A.valid) return;
    const int2 pr = pairs[p];
    if (pr.x < 0) return;
    const int i = pr.x, j = pr.y;
    const Thr th = load_threshold(A);
    const float mi = A.mu_c[i], ii = A.win[i].isig, mj = A.mu_c[j], ij = A.win[j].isig;
    for (int k = lane; k < m; k += 64) {
        zs[k] = (A.ac[i + k] - mi) * ii;
        zs[m + k] = (A.ac[j + k] - mj) * ij;
    }
    __syncthreads();
    const PlainZ<float> z{zs, zs + m};
    int t_stop = 0;
    const float res = dtw_band<float, QP>(m, A.w, 64, 1, th.dtw_ub, z, t_stop);
    if (res <= th.dtw_ub && lane == 0) {
        const float sm = sqrtf((float)m);
        const float e = pair_cell_err(A, 8.f * sm, mi, ii, mj, ij);
        push_candidate(A, i, j, res, widen_cost(sqrtf(res), e, sqrtf(2.f * (float)m), A));
    }
}

template <int QP>
__global__ __launch_bounds__(64) void k_seed_dtw(const JoinArgs *__restrict__ Ap, const int2 *pairs, int npairs) {
    extern __shared__ float zs[];
    seed_dtw_block<QP>(*Ap, pairs, npairs, (int)blockIdx.x, zs);
}

// The seed pairs of several series of a batch (one band class) in one launch.
template <int QP>
__global__ __launch_bounds__(64) void k_seed_dtw_batch(const JoinArgs *__restrict__ args,
                                                       const SeedDtwJob *__restrict__ jobs,
                                                       const int32_t *__restrict__ base, int32_t nj) {
    extern __shared__ float zs[];
    const int b = (int)blockIdx.x;
    const int k = batch_owner(base, nj, b);
    seed_dtw_block<QP>(args[jobs[k].series], jobs[k].pairs, jobs[k].npairs, b - base[k], zs);
}

}  // namespace

namespace {
// Segment-mean bound of explicit window pairs (the building block behind dtwj_paa_pairs).
__global__ __launch_bounds__(256) void k_paa_pairs(const float *__restrict__ paa, int32_t nseg, int32_t len,
                                                   const int64_t *pi, const int64_t *pj, int64_t npairs, double *out) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k < npairs) out[k] = (double)paa_lb(paa, pi[k], pj[k], nseg, (float)len);
}
}  // namespace

int launch_paa_pairs(const float *paa, int32_t nseg, int32_t len, const int64_t *pi, const int64_t *pj, int64_t npairs,
                     double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_paa_pairs, dim3((unsigned)((npairs + 255) / 256)), dim3(256), 0, s, paa, nseg, len, pi, pj,
                       npairs, out);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

namespace {
// tot: the kernel carries LB_Keogh totals with its DTW queue (several offset pairs per lane)
size_t join_lds_bytes_pw(int32_t m, bool pw, bool tot = false, bool gs = false) {
    const int q3 = dtw_queue_cap(gs);
    return sizeof(float) * (size_t)union_floats(m, pw, gs) +
           sizeof(uint32_t) * (PLCAP + keogh_queue_cap(gs) + q3 + QPCAP) + (tot ? sizeof(float) * 2 * q3 : 0);
}
// Waves per CU that LDS of the given dynamic size allows (160 KiB, allocated in 512-byte
// blocks, with the kernel's static stage counters), at most the 8 that the registers allow.
int lds_waves_per_cu(size_t bytes) {
    const size_t b = (bytes + 8 * kWcnt + 511) / 512 * 512;
    const int n = (int)(160 * 1024 / b);
    return n < 8 ? n : 8;
}
// The gs form for launches of two offset pairs per lane whose LDS then holds more waves per CU
// (at m = 600, where both forms hold 8, they measured within 0.5 % of each other).
bool join_gs(int Q, int32_t m) {
    if (Q != 2) return false;
    return lds_waves_per_cu(join_lds_bytes_pw(m, true, true, true)) >
           lds_waves_per_cu(join_lds_bytes_pw(m, true, true, false));
}
// Packed FP16 envelope staging: always for two offset pairs per lane (k_join_occ2), and for one
// pair per lane when FP32 staging of windows of length m would leave fewer than the 8 waves per
// CU its registers allow (m > 402: at m = 512, 7 waves instead of 8).
bool join_packed(int Q, int32_t m, bool tot = false) {
    if (packed_envelope(Q)) return true;
    if (Q != 1) return false;
    if (tot) return true;  // the fused-arithmetic one-pair kernels: packed staging at every m
    return 160 * 1024 / (join_lds_bytes_pw(m, false, tot) + 8 * kWcnt) < 8;
}
// The one-pair-per-lane join with the fused arithmetic (and, for bands of 16 and more, the
// suffix bound from LB_Keogh totals): launches whose widest band is 4 to 63.  The narrowest
// (w <= 3, sweep-bound) keep the separately normalised form, which measured faster there.
bool join_sw(int Q, int32_t max_w) { return Q == 1 && max_w >= 4; }
}  // namespace

size_t join_lds_bytes(int32_t m, int32_t w) {
    const int Q = dtw_qp_for(w);
    const bool tot = Q >= 2 || join_sw(Q, w);
    return join_lds_bytes_pw(m, join_packed(Q, m, tot), tot, join_gs(Q, m));
}

int launch_join_setup(JoinArgs *d_series, int32_t nseries, unsigned long long *d_work, hipStream_t s) {
    hipLaunchKernelGGL(k_join_setup, dim3((unsigned)((nseries + 63) / 64)), dim3(64), 0, s, d_series, nseries, d_work);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_join(const JoinArgs *d_series, const int64_t *d_item_base, int32_t nseries, int32_t max_m, int32_t max_w,
                unsigned long long *d_work, uint32_t *d_queue, int32_t grid, hipStream_t s) {
    const int Q = dtw_qp_for(max_w);
    const bool sw = join_sw(Q, max_w);
    const bool pw = join_packed(Q, max_m, Q >= 2 || sw);
    const size_t lds = join_lds_bytes(max_m, max_w);
    const bool gs = join_gs(Q, max_m);
    const int32_t uni = union_floats(max_m, pw, gs);
    if (lds + 8 * kWcnt > 160 * 1024) {  // + the kernel's static stage counters
        set_error("window too long for the join's LDS staging");
        return DTWJ_EINVAL;
    }
#define DTWJ_LAUNCH_JOIN(QQ, PP, SS)                                                                                  \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_join<QQ, PP, SS>), dim3(grid), dim3(64), lds, s, d_series, d_item_base, \
                       nseries, uni, d_work, d_queue)
    switch (Q) {
        case 1:
            if (sw) {
                DTWJ_LAUNCH_JOIN(1, true, true);  // join_packed() holds for these launches
            } else {
                if (pw)
                    DTWJ_LAUNCH_JOIN(1, true, false);
                else
                    DTWJ_LAUNCH_JOIN(1, false, false);
            }
            break;
        case 2:  // two waves per SIMD (packed envelope staging)
            if (gs)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_join_occ2<2, true>), dim3(grid), dim3(64), lds, s, d_series,
                                   d_item_base, nseries, uni, d_work, d_queue);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_join_occ2<2, false>), dim3(grid), dim3(64), lds, s, d_series,
                                   d_item_base, nseries, uni, d_work, d_queue);
            break;
        case 4: DTWJ_LAUNCH_JOIN(4, false, false); break;
        case 8: DTWJ_LAUNCH_JOIN(8, false, false); break;
        default: set_error("band too wide: 2w+1 must be <= 1024"); return DTWJ_EINVAL;
    }
#undef DTWJ_LAUNCH_JOIN
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_seed(const SeedArgs &S, int2 *pairs, hipStream_t s) {
    hipLaunchKernelGGL(k_seed<false>, dim3(S.n_seed + 1), dim3(64), 0, s, S, pairs);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_seed_batch(const SeedJob *jobs, const int32_t *base, int32_t nj, int32_t nblocks, hipStream_t s) {
    if (nj <= 0 || nblocks <= 0) return 0;
    hipLaunchKernelGGL(k_seed_batch, dim3(nblocks), dim3(64), 0, s, jobs, base, nj);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_seed_dtw_batch(const JoinArgs *d_args, int32_t max_m, int32_t w, const SeedDtwJob *jobs,
                          const int32_t *base, int32_t nj, int32_t nblocks, hipStream_t s) {
    if (nj <= 0 || nblocks <= 0) return 0;
    const int QP = dtw_qp_for(w);
    const size_t lds = sizeof(float) * 2 * (size_t)max_m;
    switch (QP) {
        case 1: hipLaunchKernelGGL(k_seed_dtw_batch<1>, dim3(nblocks), dim3(64), lds, s, d_args, jobs, base, nj); break;
        case 2: hipLaunchKernelGGL(k_seed_dtw_batch<2>, dim3(nblocks), dim3(64), lds, s, d_args, jobs, base, nj); break;
        case 4: hipLaunchKernelGGL(k_seed_dtw_batch<4>, dim3(nblocks), dim3(64), lds, s, d_args, jobs, base, nj); break;
        case 8: hipLaunchKernelGGL(k_seed_dtw_batch<8>, dim3(nblocks), dim3(64), lds, s, d_args, jobs, base, nj); break;
        default: set_error("band too wide: 2w+1 must be <= 1024"); return DTWJ_EINVAL;
    }
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_seed_lanes(const SeedArgs &S, int2 *pairs, hipStream_t s) {
    hipLaunchKernelGGL(k_seed<true>, dim3(S.n_seed), dim3(64), 0, s, S, pairs);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_seed_dtw(const JoinArgs *d_args, int32_t m, int32_t w, const int2 *pairs, int32_t npairs, hipStream_t s) {
    if (npairs <= 0) return 0;
    const int QP = dtw_qp_for(w);
    const size_t lds = sizeof(float) * 2 * (size_t)m;
    switch (QP) {
        case 1: hipLaunchKernelGGL(k_seed_dtw<1>, dim3(npairs), dim3(64), lds, s, d_args, pairs, npairs); break;
        case 2: hipLaunchKernelGGL(k_seed_dtw<2>, dim3(npairs), dim3(64), lds, s, d_args, pairs, npairs); break;
        case 4: hipLaunchKernelGGL(k_seed_dtw<4>, dim3(npairs), dim3(64), lds, s, d_args, pairs, npairs); break;
        case 8: hipLaunchKernelGGL(k_seed_dtw<8>, dim3(npairs), dim3(64), lds, s, d_args, pairs, npairs); break;
        default: set_error("band too wide: 2w+1 must be <= 1024"); return DTWJ_EINVAL;
    }
    DTWJ_HIP(hipGetLastError());
    return 0;
}

}  // namespace dtwj
