/*
 * CPU oracle for the top-1 banded-DTW subsequence join -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker (or the timed CPU baseline).  The product path never
 * calls it.
 *
 * Two independent oracles:
 *  (1) orc_join / orc_motif_discovery: a plain-C restatement of the reference's FP64 CPU
 *      join (CPU_version1/new_dtw_motifGUI.cpp:21-586 and the files it calls), in the same
 *      floating-point operation order, so the seed, the pruning stages, LB_KK, LB_Petitjean
 *      (CPU_version1/LB_PETITJEAN.cpp:9-123), the alpha-tuned local masks with their alpha
 *      search (CPU_version1/new_dtw_motifGUI.cpp:332-498, diag_mask_local_up / _down,
 *      compute_shared_data_local), the per-stage prune counts, the number of pairs that reach
 *      the DTW and the answer agree with the compiled reference (checked against golden vectors
 *      from oracle/_ref).
 *  (2) orc_bruteforce_top1: textbook banded DTW over every admissible pair (tiny n), with
 *      its own two-pass window statistics; shares no code with (1).
 */
#ifndef DTWJ_ORACLE_H
#define DTWJ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_join_result {
    int64_t first_min, sec_min; /* as the reference reports them (1-based when found by the
                                   sweep, 0-based seed fallback otherwise) */
    int32_t found_in_sweep;     /* 1 if the sweep updated the answer */
    int64_t i, j;               /* 0-based winning pair (i < j): the sweep's pair, or the seed
                                   pair that set the final distance when the sweep found none */
    double bsf;                 /* final distance */
    double seed_bsf;            /* best_so_far after the seed loop ("input of bsf") */
    double mp_seed;             /* mp_ed[min_index[0]], the seed before DTW refinement */
    double fast_cnt;            /* pairs pruned by diag_fast */
    double global_cnt;          /* pairs pruned by diag_mask_global / LB_KIM */
    int64_t all_cnt;            /* pairs visited */
    double kk_cnt;              /* pairs reaching LB_KK */
    double dtw_cnt;             /* pairs reaching MON_dtw (after LB_Petitjean; no local masks here) */
    double fast_stage_rounds;   /* diagonals processed */
    double global_rounds;       /* diagonals with diag_mask_global (else LB_KIM) */
    int64_t seed_i, seed_j;     /* 0-based pair that set the seed's best distance (not reported
                                   by the reference, which prints the last seed visited) */
    double pe_cnt;              /* pairs reaching LB_Petitjean (both LB_KK passed) */
    double local_rounds;        /* diagonals with diag_mask_local_up ("local_up times") */
    double local_down_rounds;   /* diagonals with diag_mask_local_down ("local_down times") */
    double local_cnt;           /* pairs pruned by the two local masks ("local prunes" x all_cnt) */
    double bad_bin;             /* local-mask count reads at bin 0, where the reference reads index
                                   -1 of a count row (outside it); read here as 0 (none on the
                                   fixtures) */
} orc_join_result;

/* CPU_version1/allmatrixop/movmean.cpp:9-33 */
void orc_mvmean(const double *a, int64_t n, int32_t m, double *mu, double *sig);
/* CPU_version1/allmatrixop/lower_upper_lemire.cpp:7-43 */
void orc_lemire(const double *a, int64_t n, int32_t w, double *lower, double *upper);
/* CPU_version1/allmatrixop/moving_mean.cpp:12-50 (compensated) */
void orc_moving_mean(const double *a, int64_t n, int32_t w, double *res);
/* CPU_version1/mpx_v2.cpp:13-150: ED matrix profile (mp: distances, mpi: 1-based index) */
int orc_mpx(const double *a, int64_t n, int32_t minlag, int32_t m, double *mp, int32_t *mpi);
/* CPU_version1/allmatrixop/min_v.cpp:40-71 (out: k+10 ints) */
void orc_min_v_k(const double *v, int64_t len, int32_t k, int32_t *out);
/* CPU_version1/MON_DTW.cpp:20-108 (EAPrunedDTW; squared cost or 1e20) */
double orc_mon_dtw(const double *lines, const double *cols, const double *cb, int32_t l, int32_t w,
                   double bsf);
/* CPU_version1/new_dtw_motifGUI.cpp:21-586 (see header comment) */
int orc_join(const double *a, int64_t n, int32_t m, int32_t w, const double *mp_ed, const int32_t *mpi,
             orc_join_result *out);
/* CPU_version1/DTWMotifDiscoveryGUI.cpp:12-27: orc_mpx then orc_join */
int orc_motif_discovery(const double *a, int64_t n, int32_t m, int32_t w, orc_join_result *out);

/* CPU_version1/compute_shared_data_global.cpp:12-46 + CPU_version1/compute_MASK_global.cpp:11-93
   (table length as CPU_version1/new_dtw_motifGUI.cpp:151-176): the global occupancy mask of
   every series position, from the caller's window statistics (mu, sig; NULL: the reference's
   movmean).  Optional: diff (n: the occupancy statistic after the coverage ramp), ths (the
   threshold, its sequential mean), len_of_table.  0 on success. */
int orc_global_mask(const double *a, int64_t n, int32_t m, int32_t w, const double *mu, const double *sig,
                    unsigned char *mask, double *diff, double *ths, int32_t *len_of_table);
/* CPU_version1/diag_mask_global.cpp:11-226 with an infinite best-so-far: the masked ball bound
   of both directions along the reference's diagonal diag (pair lo, lo + diag - 1 for lo = 0 ..
   n - m + 1 - diag; m < diag), running sums and all.  lb_fir: envelope of window lo against
   window lo + diag - 1; lb_sec: the reverse.  At lo = 0 the reference's second direction reads
   the envelope norm of window 0 (CPU_version1/diag_mask_global.cpp:79), kept as is. */
int orc_masked_ball_diag(const double *a, int64_t n, int32_t m, int32_t w, const double *mu, const double *sig,
                         int64_t diag, double *lb_fir, double *lb_sec);

/* ---- independent oracle ---- */
/* textbook Sakoe-Chiba DTW: squared cost D[m-1][m-1] */
double orc_dtw_sq(const double *x, const double *y, int32_t m, int32_t w);
/* z-normalised window with its own two-pass mean / population std */
int orc_znorm(const double *a, int64_t n, int64_t i, int32_t m, double *z);
/* every pair j - i >= m; ties -> smallest i, then smallest j; *d = sqrt(cost) */
int orc_bruteforce_top1(const double *a, int64_t n, int32_t m, int32_t w, int64_t *bi, int64_t *bj,
                        double *d);
/* textbook two-sided-envelope LB_Keogh of z_j against the radius-w envelope of z_i
   clipped at the window ends (squared) */
double orc_lb_keogh_sq(const double *zi, const double *zj, int32_t m, int32_t w);
/* LB_Keogh of window j against window i's envelope plus the LB_Petitjean increments
 * (CPU_version1/LB_PETITJEAN.cpp:9-123), no abandoning; -1 for a constant window. */
double orc_lb_petitjean_pair(const double *a, int64_t n, int32_t m, int32_t w, int64_t i, int64_t j);

#ifdef __cplusplus
}
#endif
#endif
