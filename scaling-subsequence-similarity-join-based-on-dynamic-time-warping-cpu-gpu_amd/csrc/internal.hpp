// Host-side prototypes shared by dtwj_api.cpp and the .hip translation units.
#pragma once

#include "common.hpp"
#include "../../include/dtwj.h"

namespace dtwj {

// Outputs and scratch of the per-series preprocessing (prep.hip).  All device pointers.
struct Prep {
    // inputs
    const double *a = nullptr;
    int64_t n = 0;
    int32_t m = 0, w = 0;
    // FP64 outputs
    double *U = nullptr, *L = nullptr, *v = nullptr;  // n
    double *mu = nullptr, *sig = nullptr, *muv = nullptr;  // nw
    double *c0 = nullptr;                             // 1: centring constant (min+max)/2
    // FP32 outputs
    float *ac = nullptr, *uc = nullptr, *lc = nullptr;  // n
    uint32_t *wc = nullptr;                             // n: FP16 widths uc - ac (lo), ac - lc (hi)
    float *mu_c = nullptr;                              // nw
    float *isg = nullptr;                               // nw: 1 / sigma (as in WinRow)
    WinRow *win = nullptr;                              // nw
    float *blkmax = nullptr;                            // 3 * nblk
    int32_t nblk = 0;
    float *gstat = nullptr;  // 8: max|ac|, max|mu_c|, max isig, max sig, min sig, -, -, -
    float *paa = nullptr;  // nw x 3 paa_seg segment-mean table (null: no segment-mean stage)
    int32_t paa_seg = 0;   // segments per window (paa_segments(m), 0: stage off)
    // scratch
    dd_t *P = nullptr;      // kPrepScanChannels * (n + 1) prefix sums
    dd_t *bsum = nullptr;   // kPrepScanChannels * nscan_blocks
    double *red = nullptr;  // 2 * 1024 min/max partials
    int32_t *flags = nullptr;  // [0] non-finite input, [1] constant window
};

// Channels of the compensated prefix sums: a - c0, (a - c0)^2, v - c0, (v - c0)^2, h^2, h, with
// v = (U + L) / 2 and h = (U - L) / 2.
constexpr int kPrepScanChannels = 6;
int64_t prep_scan_blocks(int64_t n);
// Largest envelope radius the preprocessing kernels take (their LDS tile holds 1024 + 2w doubles
// twice); a join's radius is at most 511.
constexpr int32_t kMaxEnvelopeRadius = 1536;
// Runs every preprocessing kernel on stream s (no host synchronisation).
int prep_series(Prep &P, hipStream_t s);

// Seed scan (join.hip): z-normalised ED on a sample of work items; pairs[k] = closest pair
// of seed item k, or (-1, -1); pairs[n_seed] = (0, m).
int launch_seed(const SeedArgs &args, int2 *pairs, hipStream_t s);
// Same scan, every lane reporting the closest pair of its four diagonals: pairs[64 k + lane].
int launch_seed_lanes(const SeedArgs &args, int2 *pairs, hipStream_t s);
// DTW of a pair list (seeds) of the series d_args (device); atomicMin into its best-so-far and
// pushes candidates like the join.
int launch_seed_dtw(const JoinArgs *d_args, int32_t m, int32_t w, const int2 *pairs, int32_t npairs, hipStream_t s);
// Batched forms: the seed scans of nj series (nblocks = sum of n_seed + 1, base = their prefix
// sums) in one launch, and the seed-pair DTWs of nj series of one band class (QP of w; LDS
// for the longest window max_m).
int launch_seed_batch(const SeedJob *jobs, const int32_t *base, int32_t nj, int32_t nblocks, hipStream_t s);
int launch_seed_dtw_batch(const JoinArgs *d_args, int32_t max_m, int32_t w, const SeedDtwJob *jobs,
                          const int32_t *base, int32_t nj, int32_t nblocks, hipStream_t s);
// Error model, validity and state resets of nseries device join descriptors; zeroes *d_work.
int launch_join_setup(JoinArgs *d_series, int32_t nseries, unsigned long long *d_work, hipStream_t s);
// The fused persistent join over nseries series (device descriptors, cumulative item counts).
int launch_join(const JoinArgs *d_series, const int64_t *d_item_base, int32_t nseries, int32_t max_m, int32_t max_w,
                unsigned long long *d_work, uint32_t *d_queue, int32_t grid, hipStream_t s);
size_t join_lds_bytes(int32_t m, int32_t w);

// DTW of an explicit pair list (dtw.hip).  prec 32: FP32 cells from ac/mu_c/isig;
// prec 64: FP64 cells from a/mu/sig.  cost: npairs doubles (device).  Reads only the
// series / statistics pointers and m, w of args (a host copy).
// Segment-mean bound (FP32, unscaled) of explicit window pairs from a table built by prep_series.
int launch_paa_pairs(const float *paa, int32_t nseg, int32_t len, const int64_t *pi, const int64_t *pj, int64_t npairs,
                     double *out, hipStream_t s);
int launch_dtw_pairs(const JoinArgs &args, const double *sig64, const int64_t *pi, const int64_t *pj,
                     int64_t npairs, int32_t prec, double *cost, hipStream_t s);

// FP64 winner of a candidate list: cost, pair, the FP32 cost the join found, and how many
// candidates were re-scored (FP32 cost <= ub).
struct Winner {
    double cost;
    int32_t i, j;
    float cost32;
    uint32_t pad;
    unsigned long long verified;
};
// Number of Winner partials launch_rescore needs in `part`.
int rescore_partials();
// Re-scores the n candidates (i, j, FP32 cost bits) of a join in FP64 (those with FP32 cost
// <= ub; cost64: n doubles of scratch) and writes the lexicographic minimum of
// (FP64 cost, i, j) to *out (device).  Reads the series / statistics pointers and m, w of A.
int launch_rescore(const JoinArgs &A, const double *sig64, const int4 *cands, int64_t n, float ub, double *cost64,
                   Winner *part, Winner *out, hipStream_t s);

// ED matrix profile (mpx.hip) with exclusion m: mp (FP64 distances) and mpi (1-based partner
// indices, 0 where no window is m or more away), nw = n - m + 1 entries each; scratch holds
// mp_scratch_doubles(n, m) doubles; flags[0] / flags[1]: non-finite sample / constant window.
int64_t mp_scratch_doubles(int64_t n, int32_t m);
size_t mp_lds_bytes(int32_t m);
int launch_matrix_profile(const double *d_a, int64_t n, int32_t m, double *scratch, int32_t *flags, double *d_mp,
                          int32_t *d_mpi, hipStream_t s);

// The reference's global occupancy mask and masked ball bound (mask.hip).  Inputs from the
// preprocessing (series, envelope, FP64 window statistics); buffers: table m x lmax counts,
// diff (n), mask (n), ths, zmax_bits (1 each), status[2] (table length, error flag: the table
// would need more than lmax bins).
struct MaskBuffers {
    const double *a = nullptr, *U = nullptr, *L = nullptr, *mu = nullptr, *sig = nullptr;
    int64_t n = 0;
    int32_t m = 0, w = 0;
    int64_t lmax = 0;
    uint32_t *table = nullptr;
    double *diff = nullptr;
    uint8_t *mask = nullptr;
    double *ths = nullptr;
    unsigned long long *zmax_bits = nullptr;
    int32_t *status = nullptr;
};
// limits of the mask kernels: the envelope radius (sparse tables of a block's span in LDS) and
// the window length (a histogram row of max(600, 200 (floor(sqrt(m - 1)) + 1)) bins in LDS)
constexpr int32_t kMaskMaxRadius = 256;
constexpr int32_t kMaskMaxWindow = 16384;
int64_t mask_table_bins(int32_t m);
size_t mask_sum_lds_bytes(int32_t m, int32_t w);
int launch_global_mask(const MaskBuffers &B, hipStream_t s);
int launch_masked_ball(const MaskBuffers &B, const int64_t *pi, const int64_t *pj, int64_t npairs, double *lb,
                       hipStream_t s);

// Fixed per-device workspace of the no-allocation entry point (noalloc.hip): its size, and the
// device address of `device`'s copy (looked up on first use, then cached).
size_t noalloc_scratch_bytes();
int noalloc_scratch(int device, void **ptr);

// LB_Keogh plus the LB_Petitjean increments of explicit pairs (dtw.hip), FP64.
int launch_lb_petitjean(const double *a, const double *mu, const double *sig, int32_t m, int32_t w, const int64_t *pi,
                        const int64_t *pj, int64_t npairs, double *out, hipStream_t s);

// Learned proposal (proposal.hip).
int mlp_param_count();
int mlp_feature_pad();
void mlp_to_params(const dtwj_mlp &M, float *P);
void params_to_mlp(const float *P, dtwj_mlp &M);
int launch_prop_features(const JoinArgs &A, const int2 *cand, int count, float *F, hipStream_t s);
// Candidate pairs sorted by (i, j) on the device (bitonic network over 64-bit keys; entries
// without a pair sort last as ~0): key holds len = pair_sort_len(count) keys, a power of two.
int64_t pair_sort_len(int64_t count);
int launch_pair_sort(const int2 *cand, int64_t count, uint64_t *key, int64_t len, hipStream_t s);
int launch_prop_pack(const float *F, int count, int rows, const float *mean, const float *scale, uint16_t *X,
                     hipStream_t s);
int launch_mlp_forward(const uint16_t *X, int count, const uint16_t *W1b, const float *P, float *out, hipStream_t s);
int launch_mlp_grad(const uint16_t *X, const float *target, const int32_t *idx, int bsz, const uint16_t *W1b,
                    const float *P, float *G, float *loss, hipStream_t s);
int launch_mlp_adam(float *P, const float *G, float *M1, float *M2, float lr, int step, uint16_t *W1b, hipStream_t s);
int launch_params_to_bf16(const float *P, uint16_t *W1b, hipStream_t s);
int launch_prop_keys(const float *F, const float *y, int count, int m, float *key, float *key_ed, hipStream_t s);

}  // namespace dtwj
