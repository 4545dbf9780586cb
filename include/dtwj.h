/*
 * dtwj.h -- C ABI of the MI355X (gfx950) top-1 subsequence similarity join under
 * Sakoe-Chiba-banded DTW.
 *
 * Problem (same as the reference join, CPU_version1/new_dtw_motifGUI.cpp:21-586 and
 * GPU_version1/src/new_dtw_motifGUI_malloc.cu:30-816):
 *   given a series a[0..n), a window length m and a band radius w, find the pair of
 *   windows (i, j), 0 <= i, j - i >= m (trivial-match exclusion, minlag = m,
 *   CPU_version1/new_dtw_motifGUI.cpp:182,364), minimising
 *       DTW_w(z_i, z_j) = sqrt(D[m-1][m-1]),
 *       D[r][c] = (z_i[r] - z_j[c])^2 + min(D[r-1][c], D[r][c-1], D[r-1][c-1]), |r - c| <= w,
 *   where z_i[k] = (a[i+k] - mu_i) / sigma_i with the population standard deviation
 *   (CPU_version1/allmatrixop/movmean.cpp:9-33, CPU_version1/new_dtw_motifGUI.cpp:41-44).
 *   Exact ties are resolved towards the smallest i, then the smallest j.
 *
 * All indices in this ABI are 0-based.  The reference prints 1-based indices
 * ("first min is i+1 sec min is j+1", CPU_version1/new_dtw_motifGUI.cpp:541-545,575);
 * the gpu_dtw command-line tool reproduces that line.
 *
 * Ownership: callers own every host buffer passed in; they are only read during the
 * call.  Device memory is owned by the library (per call, or by a dtwj_context).
 * Errors are returned as codes, never by exiting; dtwj_last_error() gives a message
 * (thread-local).
 */
#ifndef DTWJ_H
#define DTWJ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define DTWJ_OK 0
#define DTWJ_EINVAL 1    /* bad argument: m < 8, w < 0, n < 2m, NaN/Inf, constant window */
#define DTWJ_ENOMEM 2    /* host or device allocation failed */
#define DTWJ_EDEVICE 3   /* no usable HIP device, or a HIP call failed */
#define DTWJ_EIO 4       /* file could not be opened / parsed / written */
#define DTWJ_EOVERFLOW 5 /* more than 2^31 - 1 near-tie candidates (the candidate buffer is
                            regrown and the join re-run below that) */

/* ---- options and results --------------------------------------------------------- */
typedef struct dtwj_options {
    int32_t device;        /* HIP device ordinal (default 0) */
    int32_t seed_items;    /* ED-scan seed work items before the join (0 = automatic: 1024, or 4096
                              per series of a batch) */
    int32_t rows_per_item; /* rows per sweep work item, multiple of 128 (default 0 = automatic) */
    int32_t waves_per_cu;  /* persistent waves per CU (default 0 = automatic) */
    int32_t verbose;       /* print per-stage counters to stderr */
    int32_t graph;         /* batch: capture prep + seed + join in one HIP graph (default 0 = on,
                              -1 = launch kernels directly) */
    int32_t autotune;      /* batch: 1 = pick rows_per_item / waves_per_cu by timing one join of
                              the whole batch per candidate, kept per context and batch shape
                              (0 = use the fields above or defaults) */
    int32_t sweep;         /* diagonal sweep test: 0 = adaptive per work item (default), 1 = always
                              the two-sided ball bound, 2 = always the one-sided one (both exact) */
    int32_t cand_cap;      /* near-tie candidate slots of a single-series join (0 = 2^20; a batch
                              keeps 2^14 per series).  A join that pushes more is re-run with a
                              buffer sized to its count, so this only trades memory for re-runs */
    int32_t paa;           /* segment-mean (PAA) bound between LB_Kim and LB_Keogh: 0 = automatic
                              (bands w >= 8 with m >= 128; off in dtwj_top1_device_noalloc), 1 = on
                              (m >= 16), 2 = off.  It only prunes: the answer is the same */
    int32_t reserved[2];
} dtwj_options;

typedef struct dtwj_result {
    int64_t i, j;        /* 0-based winning pair, i < j, j - i >= m */
    double dist;         /* FP64-verified banded DTW distance of (i, j) */
    double dist_fp32;    /* the FP32 distance the GPU found for (i, j) */
    double seed_dist;    /* best-so-far distance entering the join (from the seed) */
    /* per-stage counters of the GPU cascade */
    int64_t pairs_total;     /* candidate pairs in the triangle: K(K+1)/2, K = n-2m+1 */
    int64_t pairs_swept;     /* pairs evaluated by the diagonal bound sweep */
    int64_t pass_ball;       /* pairs surviving the two-sided ball bound */
    int64_t pass_kim;        /* ... and the LB_Kim end-point bound */
    int64_t pass_keogh;      /* ... and the two-sided LB_Keogh bound (these reach DTW) */
    int64_t dtw_abandoned;   /* DTW evaluations stopped early by the best-so-far */
    int64_t dtw_cells;       /* DTW cells evaluated (FP32) */
    int64_t keogh_elems;     /* (pair, position) terms evaluated by LB_Keogh (both directions) */
    int64_t candidates;      /* FP32 results re-verified in FP64 */
    /* wall-clock seconds (host steady clock, device synchronised) */
    double t_total;      /* entry to result, including host<->device copies */
    double t_prep;       /* statistics, envelopes, bound tables */
    double t_seed;       /* seed scan */
    double t_join;       /* fused sweep + cascade + DTW */
    double t_verify;     /* FP64 re-verification */
    double join_kernel_ms; /* GPU time of the fused join kernel (HIP events) */
    int32_t items;       /* sweep work items (saturated at 2^31 - 1; see items_total) */
    int32_t rows_per_item; /* sweep tile height used */
    int32_t waves_per_cu;  /* persistent waves per CU used */
    int32_t reruns;      /* joins re-run because more near-tie candidates were pushed than the
                            candidate buffer held (each re-run starts from the previous run's
                            best-so-far with a buffer sized to the count; 0 normally) */
    int64_t pairs_swept_one_sided; /* of pairs_swept: tested with the one-sided ball bound */
    int64_t items_total; /* sweep work items (64-bit) */
    int64_t pass_paa;    /* of pass_kim: pairs also surviving the segment-mean bound, which go on
                            to LB_Keogh (equal to pass_kim when that stage is off) */
} dtwj_result;

/* Fills *opt with defaults.  No reference counterpart (the reference hard-codes
 * its tuning in macros, GPU_version1/alldef/GPU_parameters.h:8-18). */
void dtwj_default_options(dtwj_options *opt);

/* Work plan of a single-series top-1 join, computed on the host without touching a device:
 * the sweep's work items (claimed on the device from a 64-bit counter), tile height, diagonal
 * groups and the device bytes the join's workspace takes.  No reference counterpart (the
 * reference tiles with fixed macros, GPU_version1/alldef/GPU_parameters.h:8-18). */
typedef struct dtwj_plan_info {
    int64_t pairs_total;   /* K(K+1)/2, K = n - 2m + 1 */
    int64_t items;         /* sweep work items */
    int32_t rows_per_item; /* tile height */
    int32_t diag_groups;   /* groups of 256 diagonal offsets */
    int64_t device_bytes;  /* workspace bytes for a device of 256 compute units */
    int64_t reserved[4];
} dtwj_plan_info;
int dtwj_plan(int64_t n, int32_t m, int32_t w, const dtwj_options *opt, dtwj_plan_info *out);

/* Band radius from a ratio, the reference rule w = MAX(round(m*r), 1)
 * (GPU_version1/src/AAAA.cu:34; CPU_version1/AAAA.cpp:36). */
int32_t dtwj_band_from_ratio(int32_t m, double r);

/* Top-1 join on a host series.  Replaces the in-process entry point
 * new_dtw_motifGUI_malloc(a, m, w, mp_ed, mp_ed_index, file)
 * (GPU_version1/src/new_dtw_motifGUI_malloc.cu:30-31; CPU twin
 * CPU_version1/new_dtw_motifGUI.cpp:21-22) minus the ED seed arguments, which the
 * library computes itself on the GPU.  opt may be NULL. */
int dtwj_top1(const double *series, int64_t n, int32_t m, int32_t w, const dtwj_options *opt,
              dtwj_result *out);

/* Same, seeded with the reference's ED matrix profile (mp_ed[k] = ED distance of window
 * k to its nearest non-trivial neighbour, mp_idx_1based[k] its 1-based index), exactly the
 * extra arguments of new_dtw_motifGUI_malloc (GPU_version1/src/new_dtw_motifGUI_malloc.cu:30-31).
 * The 100 smallest entries (reference k = 100, GPU_version1/src/new_dtw_motifGUI_malloc.cu:82)
 * are verified by DTW and seed the best-so-far; the answer does not depend on the seed. */
int dtwj_top1_seeded(const double *series, int64_t n, int32_t m, int32_t w, const double *mp_ed,
                     const int32_t *mp_idx_1based, int64_t n_mp, const dtwj_options *opt,
                     dtwj_result *out);

/* One series of a batch (host memory, read during the call). */
typedef struct dtwj_series {
    const double *data;
    int64_t n;
    int32_t m;  /* window length */
    int32_t w;  /* band radius */
} dtwj_series;

/* Batched form: independent top-1 joins of `count` series of uneven lengths with mixed (m, w),
 * run as one persistent join launch fed from one device work queue (all series' work items
 * in one list), with the preprocessing, seeding and the join captured in one HIP graph.
 * out[k] / status[k] receive series k's result and status code; the call returns DTWJ_OK
 * when the batch ran (individual series may still have failed: see status), or an error that
 * concerns the whole batch.  No reference counterpart: the reference joins one series per
 * process (GPU_version1/src/AAAA.cu:17-36). */
int dtwj_top1_batch(const dtwj_series *series, int32_t count, const dtwj_options *opt, dtwj_result *out,
                    int32_t *status);

/* ---- learned proposal ------------------------------------------------------------
 * No reference counterpart (the reference has no learned component; the nearest analogue
 * is its heuristic alpha search, CPU_version1/new_dtw_motifGUI.cpp:411-468).  A 40 -> 64 -> 1
 * MLP (BF16 weights on the GPU, FP32 accumulation) predicts, from PAA-16 features of both
 * windows and cheap bounds, log(DTW distance / ED distance) of a candidate pair (<= 0: banded
 * DTW never exceeds the z-normalised Euclidean distance); candidates are ranked by
 * log ED + prediction, so an untrained (zero) network ranks by ED.  It only orders candidates:
 * the pairs it ranks first are verified by exact DTW and seed the plain join's best-so-far,
 * so the answer is the plain join's. */
#define DTWJ_MLP_IN 40
#define DTWJ_MLP_HID 64

typedef struct dtwj_mlp {
    float w1[DTWJ_MLP_HID][DTWJ_MLP_IN]; /* hidden x input */
    float b1[DTWJ_MLP_HID];
    float w2[DTWJ_MLP_HID];
    float b2;
    float feat_mean[DTWJ_MLP_IN];        /* the network sees (feature - mean) * scale */
    float feat_scale[DTWJ_MLP_IN];
} dtwj_mlp;

typedef struct dtwj_proposal_report {
    int64_t candidates;  /* pairs scored by the network */
    int32_t verified;    /* top-ranked pairs verified by exact DTW (the join's seeds) */
    int32_t reserved0;
    double seed_dist;    /* best verified distance in the network's order */
    double seed_dist_ed; /* best verified distance when as many pairs are taken in ED order */
    double regret;       /* seed_dist / final distance - 1 (0: the network found the top-1) */
    double regret_ed;    /* the same for the ED order */
    double t_propose;    /* seconds for candidates, features, network and verification */
    double reserved[3];
} dtwj_proposal_report;

/* Candidate pairs of the learned proposal and their raw features (count x DTWJ_MLP_IN,
 * row-major): the closest pair by z-normalised ED of every lane of n_items sampled sweep
 * items (64 per item).  At most cap pairs are returned. */
int dtwj_proposal_candidates(const double *series, int64_t n, int32_t m, int32_t w, int32_t n_items, int64_t cap,
                             int64_t *pi, int64_t *pj, float *features, int64_t *count);
/* Network output (predicted log(DTW / ED) distance ratio) for count raw feature rows, on the GPU. */
int dtwj_mlp_forward(const dtwj_mlp *model, const float *features, int64_t count, float *out);
/* Gradient of the mean squared error over count rows (one full batch on the GPU, the
 * kernels training uses); fills grad's w1, b1, w2, b2 and *loss. */
int dtwj_mlp_gradient(const dtwj_mlp *model, const float *features, const float *targets, int64_t count,
                      dtwj_mlp *grad, double *loss);
/* Trains model in place on the GPU (Adam, mini-batches of `batch` rows, `epochs` passes,
 * shuffled with `seed`); model->feat_mean / feat_scale must be set by the caller. */
int dtwj_mlp_train(dtwj_mlp *model, const float *features, const float *targets, int64_t count, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, double *final_loss);
/* The learned form of the top-1 join: candidates from n_items sampled items are scored by
 * the network, the n_verify best-ranked are verified by exact DTW and seed the plain join. */
int dtwj_top1_learned(const double *series, int64_t n, int32_t m, int32_t w, const dtwj_mlp *model,
                      int32_t n_items, int32_t n_verify, const dtwj_options *opt, dtwj_result *out,
                      dtwj_proposal_report *report);

/* Reusable device context: keeps device buffers between calls (for repeated joins). */
typedef struct dtwj_context dtwj_context;
int dtwj_context_create(int32_t device, dtwj_context **ctx);
void dtwj_context_destroy(dtwj_context *ctx);
/* Top-1 join on a series already resident in device memory (d_series: n doubles on the
 * context's device).  Used to time the join with inputs resident in HBM. */
int dtwj_top1_device(dtwj_context *ctx, const double *d_series, int64_t n, int32_t m, int32_t w,
                     const dtwj_options *opt, dtwj_result *out);
/* The same join with no device allocation, for callers that cannot allocate (code that runs
 * under HIP graph capture, tight serving loops, harnesses that forbid allocation).  No
 * counterpart in the reference.  Every buffer the join needs beyond its input and output
 * lives in one fixed 64 MiB workspace per GPU (a device array of the library, held until the
 * process ends), which limits the series to about 2^17 samples (n <= 131072 at any m, w;
 * larger inputs return DTWJ_EINVAL with the bytes needed).  d_series: n doubles on the device
 * opt->device (default 0); stream: a hipStream_t to run on (NULL = the null stream).  Calls no
 * hipMalloc / hipFree.  Calls on one GPU share its workspace: calls that overlap in time on one
 * GPU race and must be serialised by the caller.  A series with more near-tie candidates than
 * the workspace's 2^16 slots returns DTWJ_EOVERFLOW (dtwj_top1_device regrows instead).  The
 * join time is taken on the host clock (no events are created). */
int dtwj_top1_device_noalloc(const double *d_series, int64_t n, int32_t m, int32_t w, const dtwj_options *opt,
                             void *stream, dtwj_result *out);
/* Copies n doubles to a device buffer owned by the context (valid until the next call of
 * this function or the context's destruction) and returns its device address. */
int dtwj_context_upload(dtwj_context *ctx, const double *series, int64_t n, const double **d_series);
/* dtwj_top1_batch on a context: device memory and the autotuner's choices are kept between
 * calls (no reference counterpart). */
int dtwj_context_top1_batch(dtwj_context *ctx, const dtwj_series *series, int32_t count, const dtwj_options *opt,
                            dtwj_result *out, int32_t *status);

/* ED matrix profile with trivial-match exclusion m, on the GPU in FP64.  Replaces mpx_v2(TS,
 * minlag = m, m) (GPU_version1/src/mpx_v2.cu:14-150, CPU_version1/mpx_v2.cpp:13-150) as
 * DTWMotifDiscoveryGUI calls it (GPU_version1/src/DTWMotifDiscoveryGUI.cu:20-24): mp_ed[t] =
 * z-normalised Euclidean distance of window t to its nearest window p with |t - p| >= m,
 * mp_idx_1based[t] = p + 1 (0 and sqrt(4m) where no window is that far).  n - m + 1 entries each.
 * Partners whose squared distances agree to 2^-21 relative count as tied: the smallest lag |t - p|
 * wins, the left one at equal lags (the reference's visiting order); the distance is the chosen
 * pair's own FP64 value.  Non-finite samples and constant windows return DTWJ_EINVAL. */
int dtwj_matrix_profile(const double *series, int64_t n, int32_t m, double *mp_ed, int32_t *mp_idx_1based);

/* Whole pipeline of the reference CLI for one (m, w): DTWMotifDiscoveryGUI(TS, m, w, file)
 * (GPU_version1/src/DTWMotifDiscoveryGUI.cu:12-28): the ED matrix profile (dtwj_matrix_profile),
 * the reference's seed (its 100 smallest entries in scan order, GPU_version1/allmatrixop/min_v.cu:15-46,
 * each pair verified by FP64 DTW, GPU_version1/src/new_dtw_motifGUI_malloc.cu:82-107), then the
 * join seeded with those pairs; prints the reference's result lines to stdout ("input of bsf"
 * is that seed) and appends one "%lf\n" seconds line to time_file
 * (GPU_version1/src/new_dtw_motifGUI_malloc.cu:744-753).  time_file may be NULL.  out->seed_dist
 * is the FP64 seed; out->t_total covers the join only, after the profile (the reference's timer
 * also starts after its profile, GPU_version1/src/new_dtw_motifGUI_malloc.cu:602). */
int dtwj_motif_discovery(const double *series, int64_t n, int32_t m, int32_t w, const char *time_file,
                         const dtwj_options *opt, dtwj_result *out);

/* Text loader: whitespace-separated decimal numbers -> malloc'ed double array (free with
 * dtwj_free).  Replaces loadfile (GPU_version1/src/loadfile.cu:14-33), which reads FP32 on the
 * GPU build and exits with status 0 on an unreadable file; this returns DTWJ_EIO instead. */
int dtwj_load_text(const char *path, double **out, int64_t *n);
void dtwj_free(void *p);

/* Message of the last error on this thread ("" if none). */
const char *dtwj_last_error(void);

/* Number of visible HIP devices (0 when none; never fails). */
int32_t dtwj_device_count(void);

/* ---- building blocks exposed for tests and tools (host pointers in and out) -------- */

/* Sliding mean and population standard deviation of every length-m window, computed on
 * the GPU from compensated (double-double) FP64 prefix sums.  Reference: mvmean
 * (CPU_version1/allmatrixop/movmean.cpp:9-33).  mu, sigma: n-m+1 doubles each. */
int dtwj_sliding_stats(const double *series, int64_t n, int32_t m, double *mu, double *sigma);

/* Lower/upper envelope of radius w clipped at the series ends, computed on the GPU.
 * Reference: lower_upper_lemire (CPU_version1/allmatrixop/lower_upper_lemire.cpp:7-43).
 * min(w, n - 1) must be <= 1536 (a radius of n - 1 or more spans the whole series), else
 * DTWJ_EINVAL. */
int dtwj_envelope(const double *series, int64_t n, int32_t w, double *lower, double *upper);

/* Banded DTW (squared cost D[m-1][m-1], not its square root) of z-normalised windows for
 * an explicit pair list, FP32 cells on the GPU (prec = 32) or FP64 cells (prec = 64).
 * Reference kernels: DTW_stairs / DTW_stairs_for_block (GPU_version1/src/DTW_stairs.cu:8-599). */
int dtwj_dtw_pairs(const double *series, int64_t n, int32_t m, int32_t w, const int64_t *pi,
                   const int64_t *pj, int64_t npairs, int32_t prec, double *cost);

/* LB_Keogh of window j against the radius-w envelope of window i (all m positions) plus the
 * reference's LB_Petitjean increments on positions [3, m - 3): window j projected into window
 * i's envelope, the radius-w envelope of that projection, and what window i pays outside it
 * beyond window j's own envelope.  FP64 on z-normalised windows, no abandoning; m <= 4096.
 * Replaces lbPetitjean_new (CPU_version1/LB_PETITJEAN.cpp:9-123) as the reference calls it in
 * its CPU join (CPU_version1/new_dtw_motifGUI.cpp:514-520, q = window j, x = window i); the
 * join's own cascade does not run it (a building block, like the masked ball bound). */
int dtwj_lb_petitjean_pairs(const double *series, int64_t n, int32_t m, int32_t w, const int64_t *pi,
                            const int64_t *pj, int64_t npairs, double *lb);

/* Segment-mean (PAA) lower bound of LB_Keogh that the join applies between LB_Kim and
 * LB_Keogh, for an explicit pair list: each window cut into S segments of len = floor(m / S)
 * samples (S = 32 for m >= 512, else 16), len * sum over segments of max(Z - U, L - Z, 0)^2 of
 * one window's segment means
 * against the other's envelope segment means (raw-series envelope of radius w), the larger of
 * the two directions, as the kernel computes it in FP32 (before its safety scaling).  m >= 16.
 * No reference counterpart (the reference's cascade has no such stage; it stands for the
 * LB_Keogh work of CPU_version1/lb_keogh.cpp:9-72 that it saves). */
int dtwj_paa_pairs(const double *series, int64_t n, int32_t m, int32_t w, const int64_t *pi,
                   const int64_t *pj, int64_t npairs, double *lb);

/* The reference's global occupancy mask of the series' positions, replacing
 * compute_shared_data_global + compute_MASK_global (CPU_version1/compute_shared_data_global.cpp:12-46,
 * CPU_version1/compute_MASK_global.cpp:11-93; GPU_version1/src/compute_shared_data_global.cu:48-81,
 * GPU_version1/src/compute_MASK_global.cu:92-171; table length as
 * CPU_version1/new_dtw_motifGUI.cpp:151-176), computed on the GPU from the library's FP64 window
 * statistics with the reference's operations in its order: per window offset a histogram of
 * the z-values in bins of 0.01, per (window, offset) the number of windows inside that window's
 * clipped envelope there, summed per position over its windows and divided by its coverage,
 * MASK = 1 where that is at most its mean.  mask: n bytes (0 / 1).  Optional (may be NULL):
 * occupancy (n doubles, the statistic compared), threshold (its mean), len_of_table (histogram
 * bins per offset).  Needs 8 <= m <= 16384 and w <= 256.  The join does not use the mask: the
 * masked ball bound's running sums would cost its sweep more per pair than the cascade work
 * they remove (the cascade's own bounds after the ball bound are cheap on the GPU). */
int dtwj_global_mask(const double *series, int64_t n, int32_t m, int32_t w, uint8_t *mask,
                     double *occupancy, double *threshold, int32_t *len_of_table);

/* The reference's masked ball bound (CPU_version1/diag_mask_global.cpp:11-226,
 * GPU_version1/lib/diag_mask_global.cu:12-230) for an explicit pair list, both directions:
 * lb[2p] with the envelope of window pi[p] against window pj[p], lb[2p + 1] the reverse.  Over the
 * offsets K = {k in [3, m-4] : MASK[e + k] = 1} of the envelope window e (MASK as
 * dtwj_global_mask), dist = sum_K (z_d - Uhat_e)^2 + (z_d - Lhat_e)^2 and DUL2 = the squared norm
 * of (U - L) / sigma_e over the masked positions of the whole window; lb = 0 if dist < DUL2, else
 * (sqrt(2 dist - DUL2) - sqrt(DUL2)) / 2: a lower bound of the banded DTW distance (not squared).
 * FP64, direct sums (the reference keeps running sums along a diagonal).  Same limits. */
int dtwj_masked_ball_pairs(const double *series, int64_t n, int32_t m, int32_t w, const int64_t *pi,
                           const int64_t *pj, int64_t npairs, double *lb);

/* ---- synthetic workloads ------------------------------------------------------------ */
typedef struct dtwj_gen_spec {
    int64_t n;             /* series length */
    uint64_t seed;         /* RNG seed */
    double step_std;       /* random-walk step standard deviation */
    double step_std_alt;   /* alternate step std for regime shifts (used when regime_len > 0) */
    int64_t regime_len;    /* samples per regime (0 = no regime shifts) */
    int32_t n_plants;      /* planted motif pairs */
    int32_t plant_len;     /* planted window length */
    double warp;           /* time-warp ratio of the copy (0.05 = +-5 %) */
    double noise_std;      /* std of the noise added to the copy */
    const int64_t *plant_src; /* optional explicit source offsets (n_plants), NULL = random */
    const int64_t *plant_dst; /* optional explicit destination offsets */
} dtwj_gen_spec;

void dtwj_gen_spec_default(dtwj_gen_spec *spec);
/* Writes spec->n doubles to out; src_out/dst_out (may be NULL) receive the plant offsets. */
int dtwj_generate(const dtwj_gen_spec *spec, double *out, int64_t *src_out, int64_t *dst_out);

#ifdef __cplusplus
}
#endif

#endif /* DTWJ_H */
