/*
 * CPU oracle for the top-1 banded-DTW subsequence join -- TEST INFRASTRUCTURE ONLY.
 * See dtwj_oracle.h for what it restates and what it leaves out.  Every function cites
 * the reference file:line it follows; expressions keep the reference's evaluation order
 * (compile without -ffast-math and without FMA contraction) so that results agree bit for
 * bit with the compiled reference CPU path.
 */
#include "dtwj_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INF 1E20 /* POSITIVE_INFINITY of CPU_version1/MON_DTW.cpp:9 */
#define ORC_MIN(x, y) ((x) < (y) ? (x) : (y))
#define ORC_MAX(x, y) ((x) > (y) ? (x) : (y))
#define ORC_DIST(x, y) (((x) - (y)) * ((x) - (y)))
/* std::min(a, b) returns a unless b < a */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------------------ */
/* CPU_version1/allmatrixop/movmean.cpp:9-33 -- running (uncompensated) sums */
void orc_mvmean(const double *a, int64_t n, int32_t l, double *miu, double *si) {
    double sum1 = 0, sum2 = 0;
    for (int i = 0; i < l; i++) {
        sum1 += a[i];
        sum2 += a[i] * a[i];
    }
    long long ll = n - l + 1;
    for (long long i = 0; i < ll - 1; i++) {
        miu[i] = sum1 / l;
        si[i] = sqrt(sum2 / l - miu[i] * miu[i]);
        sum1 = sum1 - a[i] + a[i + l];
        sum2 = sum2 - a[i] * a[i] + a[i + l] * a[i + l];
    }
    miu[ll - 1] = sum1 / l;
    si[ll - 1] = sqrt(sum2 / l - miu[ll - 1] * miu[ll - 1]);
}

/* CPU_version1/allmatrixop/lower_upper_lemire.cpp:7-43 -- monotone-deque min/max of
 * radius r, clipped at both ends */
void orc_lemire(const double *a, int64_t n, int32_t r, double *l, double *u) {
    int64_t k = 2 * (int64_t)r + 1;
    int64_t *q_u = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t *q_l = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t head_l = 0, tail_l = -1, head_u = 0, tail_u = -1, li = 0, ui = 0;
    for (int64_t i = 0; i < k - 1 - r && i < n; i++) {
        while (head_l <= tail_l && a[q_l[tail_l]] >= a[i]) tail_l--;
        q_l[++tail_l] = i;
        while (head_u <= tail_u && a[q_u[tail_u]] <= a[i]) tail_u--;
        q_u[++tail_u] = i;
    }
    for (int64_t i = k - 1 - r; i < n; i++) {
        while (head_l <= tail_l && a[q_l[tail_l]] >= a[i]) tail_l--;
        q_l[++tail_l] = i;
        while (q_l[head_l] <= i - k) head_l++;
        l[li++] = a[q_l[head_l]];
        while (head_u <= tail_u && a[q_u[tail_u]] <= a[i]) tail_u--;
        q_u[++tail_u] = i;
        while (q_u[head_u] <= i - k) head_u++;
        u[ui++] = a[q_u[head_u]];
    }
    for (int64_t i = n; i < n + r; i++) {
        while (q_u[head_u] <= i - k) head_u++;
        u[ui++] = a[q_u[head_u]];
        while (q_l[head_l] <= i - k) head_l++;
        l[li++] = a[q_l[head_l]];
    }
    free(q_u);
    free(q_l);
}

/* CPU_version1/allmatrixop/moving_mean.cpp:12-50 -- compensated moving mean, length n-w+1 */
void orc_moving_mean(const double *a, int64_t n, int32_t w, double *res) {
    double p = a[0], s = 0, x, z;
    for (int i = 2; i <= w; i++) {
        x = p + a[i - 1];
        z = x - p;
        s += (p - x + z) + a[i - 1] - z;
        p = x;
    }
    res[0] = p + s;
    for (int64_t i = w + 1; i <= n; i++) {
        x = p - a[i - w - 1];
        z = x - p;
        s += p - x + z - a[i - w - 1] - z;
        p = x;
        x = p + a[i - 1];
        z = x - p;
        s += p - x + z + a[i - 1] - z;
        p = x;
        res[i - w] = p + s;
    }
    for (int64_t i = 1; i <= n - w + 1; i++) res[i - 1] = res[i - 1] / w;
}

/* CPU_version1/allmatrixop/norm_vector.cpp (b == 2) */
static double norm2(const double *v, int64_t len) {
    double sum = 0;
    for (int64_t i = 0; i < len; i++) sum += pow(v[i], 2);
    return pow(sum, 0.5);
}

/* CPU_version1/mpx_v2.cpp:13-150 -- ED matrix profile with exclusion minlag (the motif and
 * discord searches at :116-120 feed nothing downstream and are not restated).  Non-finite
 * input is rejected instead of the reference's NaN masking (:19-58). */
int orc_mpx(const double *ts, int64_t n, int32_t minlag, int32_t m, double *mp_out, int32_t *mpi) {
    int64_t subcount = n - m + 1;
    for (int64_t t = 0; t < n; t++)
        if (!isfinite(ts[t])) return 1;
    double *mu = (double *)malloc(sizeof(double) * subcount);
    double *mus = (double *)malloc(sizeof(double) * (n - m + 2));
    double *invnorm = (double *)malloc(sizeof(double) * subcount);
    double *tmp = (double *)malloc(sizeof(double) * m);
    double *dr_bwd = (double *)malloc(sizeof(double) * subcount);
    double *dc_bwd = (double *)malloc(sizeof(double) * subcount);
    double *dr_fwd = (double *)malloc(sizeof(double) * subcount);
    double *dc_fwd = (double *)malloc(sizeof(double) * subcount);
    double *mp = (double *)malloc(sizeof(double) * subcount);
    orc_moving_mean(ts, n, m, mu);
    orc_moving_mean(ts, n, m - 1, mus);
    for (int64_t i = 1; i <= subcount; i++) {
        for (int64_t j = i; j <= i + m - 1; j++) tmp[j - i] = ts[j - 1] - mu[i - 1];
        invnorm[i - 1] = 1 / norm2(tmp, m);
    }
    for (int64_t i = 0; i < subcount; i++)
        if (!isfinite(invnorm[i])) invnorm[i] = NAN;
    dr_bwd[0] = 0;
    dc_bwd[0] = 0;
    for (int64_t k = 1; k < subcount; k++) {
        dr_bwd[k] = ts[k - 1] - mu[k - 1];
        dc_bwd[k] = ts[k - 1] - mus[k];
    }
    for (int64_t k = 0; k < subcount; k++) {
        dr_fwd[k] = ts[m - 1 + k] - mu[k];
        dc_fwd[k] = ts[m - 1 + k] - mus[k];
    }
    for (int64_t k = 0; k < subcount; k++) {
        mp[k] = -1;
        mpi[k] = 0;
    }
    for (int64_t diag = minlag + 1; diag <= subcount; diag++) {
        double cov_ = 0;
        for (int64_t k = 0; k < m; k++) cov_ += (ts[diag - 1 + k] - mu[diag - 1]) * (ts[k] - mu[0]);
        for (int64_t row = 1; row <= subcount - diag + 1; row++) {
            int64_t col = diag + row - 1;
            if (row > 1) cov_ = cov_ - dr_bwd[row - 1] * dc_bwd[col - 1] + dr_fwd[row - 1] * dc_fwd[col - 1];
            double corr_ = cov_ * invnorm[row - 1] * invnorm[col - 1];
            if (corr_ > mp[row - 1]) {
                mp[row - 1] = corr_;
                mpi[row - 1] = (int32_t)col;
            }
            if (corr_ > mp[col - 1]) {
                mp[col - 1] = corr_;
                mpi[col - 1] = (int32_t)row;
            }
        }
    }
    for (int64_t k = 0; k < subcount; k++) {
        double v = 2 * m * (1 - mp[k]);
        mp_out[k] = isnan(v) ? NAN : sqrt(ORC_MAX(0, v));
    }
    free(mu); free(mus); free(invnorm); free(tmp);
    free(dr_bwd); free(dc_bwd); free(dr_fwd); free(dc_fwd); free(mp);
    return 0;
}

/* CPU_version1/allmatrixop/min_v.cpp:40-71 -- indices strictly below the k-th smallest
 * value, in scan order, zero-padded to k+10 (a max-heap of the k smallest) */
void orc_min_v_k(const double *v, int64_t len, int32_t k, int32_t *out) {
    double *heap = (double *)malloc(sizeof(double) * k);
    for (int i = 0; i < k; i++) heap[i] = INFINITY;
    /* binary max-heap; all INFINITY is a valid heap */
    for (int64_t i = 0; i < len; i++) {
        if (v[i] < heap[0]) {
            heap[0] = v[i]; /* pop max, push v[i]: replace root and sift down */
            int p = 0;
            for (;;) {
                int c = 2 * p + 1;
                if (c >= k) break;
                if (c + 1 < k && heap[c + 1] > heap[c]) c++;
                if (heap[c] > heap[p]) {
                    double t = heap[c];
                    heap[c] = heap[p];
                    heap[p] = t;
                    p = c;
                } else
                    break;
            }
        }
    }
    double topK = heap[0];
    for (int i = 0; i < k + 10; i++) out[i] = 0;
    int idx = 0;
    for (int64_t i = 0; i < len; i++)
        if (v[i] < topK && idx < k + 10) out[idx++] = (int32_t)i;
    free(heap);
}

/* CPU_version1/lb_keogh.cpp:74-146 */
static int lb_kk_first(const double *q, const double *t, const double *U, const double *L, int64_t seqlen,
                       double threshold_2, double *cb) {
    double dist = 0.0, temp_u, temp_l;
    int m = (int)seqlen;
    double d = 0;
    double x0 = t[0];
    double y0 = t[(m - 1)];
    double dleft = ORC_DIST(x0, q[0]);
    double dright = ORC_DIST(y0, q[m - 1]);
    double x1 = (t[(1)]);
    d = smin(ORC_DIST(x1, q[0]), ORC_DIST(x0, q[1]));
    d = smin(d, ORC_DIST(x1, q[1]));
    dleft += d;
    double y1 = (t[(m - 2)]);
    d = smin(ORC_DIST(y1, q[m - 1]), ORC_DIST(y0, q[m - 2]));
    d = smin(d, ORC_DIST(y1, q[m - 2]));
    dright += d;
    if (dleft + dright >= threshold_2) {
        return 0;
    } else {
        double x2 = (t[(2)]);
        d = smin(ORC_DIST(x0, q[2]), ORC_DIST(x1, q[2]));
        d = smin(d, ORC_DIST(x2, q[2]));
        d = smin(d, ORC_DIST(x2, q[1]));
        d = smin(d, ORC_DIST(x2, q[0]));
        dleft += d;
        double y2 = (t[(m - 3)]);
        d = smin(ORC_DIST(y0, q[m - 3]), ORC_DIST(y1, q[m - 3]));
        d = smin(d, ORC_DIST(y2, q[m - 3]));
        d = smin(d, ORC_DIST(y2, q[m - 2]));
        d = smin(d, ORC_DIST(y2, q[m - 1]));
        dright += d;
        if (dleft + dright > threshold_2) return 0;
    }
    cb[0] = dleft;
    cb[1] = 0;
    cb[2] = 0;
    cb[seqlen - 3] = dright;
    cb[seqlen - 2] = 0;
    cb[seqlen - 1] = 0;
    dist = dleft + dright;
    for (int64_t i = 3; i < seqlen - 3; i++) {
        d = 0;
        temp_u = U[i];
        temp_l = L[i];
        if (q[i] > temp_u) {
            d = (q[i] - temp_u) * (q[i] - temp_u);
        } else if (q[i] < temp_l) {
            d = (q[i] - temp_l) * (q[i] - temp_l);
        }
        dist += d;
        if (threshold_2 < dist) return 0;
        cb[i] = d;
    }
    return 1;
}

/* CPU_version1/lb_keogh.cpp:9-72 */
static int lb_kk(const double *q, const double *U, const double *L, int64_t seqlen, double threshold_2,
                 double *cb, const double *special, double *lbk) {
    double dist = 0.0, temp_u, temp_l;
    dist = special[0] + special[1];
    double temp_dist = 0;
    for (int64_t i = 0; i < 3; i++) {
        double d = 0;
        temp_u = U[i];
        temp_l = L[i];
        if (q[i] > temp_u) {
            d = (q[i] - temp_u) * (q[i] - temp_u);
        } else if (q[i] < temp_l) {
            d = (q[i] - temp_l) * (q[i] - temp_l);
        }
        temp_dist += d;
    }
    for (int64_t i = seqlen - 3; i < seqlen; i++) {
        double d = 0;
        temp_u = U[i];
        temp_l = L[i];
        if (q[i] > temp_u) {
            d = (q[i] - temp_u) * (q[i] - temp_u);
        } else if (q[i] < temp_l) {
            d = (q[i] - temp_l) * (q[i] - temp_l);
        }
        temp_dist += d;
    }
    cb[0] = special[0];
    cb[1] = 0;
    cb[2] = 0;
    cb[seqlen - 3] = special[1];
    cb[seqlen - 2] = 0;
    cb[seqlen - 1] = 0;
    dist = ORC_MAX(temp_dist, dist);
    for (int64_t i = 3; i < seqlen - 3; i++) {
        double d = 0;
        temp_u = U[i];
        temp_l = L[i];
        if (q[i] > temp_u) {
            d = (q[i] - temp_u) * (q[i] - temp_u);
        } else if (q[i] < temp_l) {
            d = (q[i] - temp_l) * (q[i] - temp_l);
        }
        dist += d;
        if (threshold_2 < dist) return 0;
        cb[i] = d;
    }
    *lbk = dist;
    return 1;
}

/* LB_Petitjean, CPU_version1/LB_PETITJEAN.cpp:9-123 (lbPetitjean_new).  q is the window
 * LB_Keogh was taken of (uq, lq its per-window envelope), x the other window (ux, lx): the
 * projection p of q (q clamped into [lx, ux] on positions [3, m-3), q elsewhere), its clipped
 * radius-w envelope (up, lp: the reference's monotone deques compute the same clipped sliding
 * max / min as orc_lemire), then for positions [3, m-3) the extra cost of x outside that
 * envelope -- reduced by what the part beyond q's own envelope already paid -- is added to lbk
 * (which holds LB_Keogh of q against x's envelope) and to cb, abandoning once the sum exceeds
 * threshold_2.  The reference interleaves the three passes with a lag of w; each value is
 * computed from the same operands in the same order here. */
static int lb_petitjean(double *p, double *up, double *lp, const double *q, const double *uq, const double *lq,
                        const double *x, const double *ux, const double *lx, int w, int m, double threshold_2,
                        double *lbk, double *cb) {
    memcpy(p, q, sizeof(double) * (size_t)m);
    for (int i = 3; i < m - 3; i++) {
        if (q[i] > ux[i])
            p[i] = ux[i];
        else if (q[i] < lx[i])
            p[i] = lx[i];
    }
    orc_lemire(p, m, w, lp, up);
    double d = *lbk;
    for (int ii = 3; ii < m - 3; ii++) {
        double local_dist = 0;
        if (x[ii] > up[ii] && up[ii] > uq[ii])
            local_dist = ORC_DIST(x[ii], uq[ii]) - ORC_DIST(up[ii], uq[ii]);
        else if (x[ii] < lp[ii] && lp[ii] < lq[ii])
            local_dist = ORC_DIST(x[ii], lq[ii]) - ORC_DIST(lp[ii], lq[ii]);
        else if (x[ii] > up[ii])
            local_dist = ORC_DIST(x[ii], up[ii]);
        else if (x[ii] < lp[ii])
            local_dist = ORC_DIST(x[ii], lp[ii]);
        cb[ii] += local_dist;
        d += local_dist;
        if (d > threshold_2) return 0;
    }
    *lbk = d;
    return 1;
}

/* CPU_version1/MON_DTW.cpp:11-18 */
static inline size_t cap_start(size_t index, size_t window) { return index > window ? index - window : 0; }
static inline size_t cap_stop(size_t index, size_t window, size_t end) {
    if (window < end && index + 1 < end - window) return index + window + 1;
    return end;
}

/* CPU_version1/MON_DTW.cpp:20-108 -- EAPrunedDTW: exact squared cost or 1e20 */
double orc_mon_dtw(const double *lines, const double *cols, const double *cb, int32_t l, int32_t w_in,
                   double bsf) {
    size_t w = (size_t)w_in;
    double UB = bsf - cb[w + 1];
    const size_t nbcols = (size_t)l;
    const size_t nblines = (size_t)l;
    if (w > nblines) w = nblines;
    double *buffers = (double *)malloc(sizeof(double) * (1 + nbcols) * 2);
    for (size_t k = 0; k < (1 + nbcols) * 2; k++) buffers[k] = ORC_INF;
    size_t c = 1, p = nbcols + 2;
    buffers[c - 1] = 0;
    size_t next_start = 0;
    size_t pruning_point = 0;
    double result;
    for (size_t i = 0; i < nblines; ++i) {
        size_t tmp = c;
        c = p;
        p = tmp;
        if (i + w < nbcols) UB = bsf - cb[i + w + 1];
        const double li = lines[i];
        const size_t jStop = cap_stop(i, w, nbcols);
        const size_t jStart = ORC_MAX(cap_start(i, w), next_start);
        size_t next_pruning_point = jStart;
        size_t j = jStart;
        next_start = jStart;
        buffers[c + j - 1] = ORC_INF;
        double cost = ORC_INF;
        for (; j == next_start && j < pruning_point; ++j) {
            const double d = ORC_DIST(li, cols[j]);
            cost = smin(buffers[p + j - 1], buffers[p + j]) + d;
            buffers[c + j] = cost;
            if (cost <= UB) {
                next_pruning_point = j + 1;
            } else {
                ++next_start;
            }
        }
        for (; j < pruning_point; ++j) {
            const double d = ORC_DIST(li, cols[j]);
            cost = smin(cost, smin(buffers[p + j - 1], buffers[p + j])) + d;
            buffers[c + j] = cost;
            if (cost <= UB) next_pruning_point = j + 1;
        }
        if (j < jStop) {
            const double d = ORC_DIST(li, cols[j]);
            if (j == next_start) {
                cost = buffers[p + j - 1] + d;
                buffers[c + j] = cost;
                if (cost <= UB) {
                    next_pruning_point = j + 1;
                } else {
                    free(buffers);
                    return ORC_INF;
                }
            } else {
                cost = smin(cost, buffers[p + j - 1]) + d;
                buffers[c + j] = cost;
                if (cost <= UB) next_pruning_point = j + 1;
            }
            ++j;
        } else if (j == next_start) {
            free(buffers);
            return ORC_INF;
        }
        for (; j == next_pruning_point && j < jStop; ++j) {
            const double d = ORC_DIST(li, cols[j]);
            cost = cost + d;
            buffers[c + j] = cost;
            if (cost <= UB) ++next_pruning_point;
        }
        pruning_point = next_pruning_point;
    }
    if (pruning_point != nbcols)
        result = ORC_INF;
    else
        result = buffers[c + nbcols - 1];
    free(buffers);
    return result;
}

/* ------------------------------------------------------------------------------------ */
/* state shared by the join's stage functions */
typedef struct {
    int64_t len, subcount;
    int m, w;
    const double *a;
    double *mu, *sig, *invsig;
    double *subs, *subsU, *subsL; /* subcount x m each */
    double *UTS, *LTS;
    /* diag_fast tables */
    double *dr_bwd, *dc_bwd, *dr_fwd, *dc_fwd, *sumU_sumL, *normUL_trans, *DUL2_fast, *DUL_fast;
    /* diag_mask_global tables */
    double *UTS_g, *LTS_g, *MASK, *TS2;
    double *dr_bwd_g, *dr_fwd_g, *dc_bwd_g, *dc_fwd_g, *sumMASK_g, *sumU_sumL_g, *normUL_g, *DUL2_g, *DUL_g;
    double *dr_bwdM, *dr_fwdM, *dc_bwdT2, *dc_fwdT2;
    double *special; /* subcount x 2 */
    unsigned char *lb_vector;
    double *lb_vector_new;
} join_state;

#define SUB(st, r) ((st)->subs + (int64_t)(r) * (st)->m)
#define SUBU(st, r) ((st)->subsU + (int64_t)(r) * (st)->m)
#define SUBL(st, r) ((st)->subsL + (int64_t)(r) * (st)->m)

/* CPU_version1/allmatrixop/movsum.cpp:6-27 (window b+1, output length len-b) */
static void movsum(const double *ts, int64_t ts_size, int b, double *result) {
    int64_t len = ts_size - b;
    double sum = 0;
    for (int i = 0; i <= b; i++) sum += ts[i];
    for (int64_t i = 0; i < len - 1; i++) {
        result[i] = sum;
        sum = sum - ts[i] + ts[i + b + 1];
    }
    result[len - 1] = sum;
}

/* CPU_version1/diag_fast.cpp:12-88 -- two-sided ball bound along one diagonal */
static void diag_fast(join_state *st, int diag, double *cnt, double bsf) {
    const double *ts = st->a;
    int subseqlen = st->m;
    int64_t subcount = st->subcount;
    const double *mu = st->mu, *invsig = st->invsig, *sumU_sumL = st->sumU_sumL;
    const double *nt = st->normUL_trans, *DUL = st->DUL_fast, *DUL2 = st->DUL2_fast;
    double lb, fir = 0, sec = 0;
    for (int k = 0; k < subseqlen; k++) fir += (ts + diag - 1)[k] * (st->UTS[k] + st->LTS[k]);
    for (int k = 0; k < subseqlen; k++) sec += ts[k] * ((st->UTS + diag - 1)[k] + (st->LTS + diag - 1)[k]);
    int64_t row = 0, col = diag - 1;
    double tt = (fir - mu[col] * sumU_sumL[row]) * invsig[col];
    double local_del = nt[row] - 2 * tt;
    lb = 0.5 * (sqrt(2 * ((local_del * invsig[row] + 2 * subseqlen * mu[row] * mu[row] * invsig[row] * invsig[row]) +
                          2 * subseqlen) -
                     DUL2[row]) -
                DUL[row]);
    if (lb > bsf) {
        (*cnt)++;
        st->lb_vector[row] = 1;
    } else {
        col = 0;
        row = diag - 1;
        tt = (sec - mu[col] * sumU_sumL[row]) * invsig[col];
        local_del = nt[row] - 2 * tt;
        lb = 0.5 * (sqrt(2 * ((local_del + 2 * subseqlen * mu[row] * mu[row] * invsig[row]) * invsig[row] +
                              2 * subseqlen) -
                         DUL2[row]) -
                    DUL[row]);
        if (lb > bsf) {
            st->lb_vector[col] = 1;
            (*cnt)++;
        }
    }
    for (int64_t lo = 1; lo < subcount - diag + 1; lo++) {
        int64_t hi = diag + lo - 1;
        fir = fir - st->dr_bwd[lo] * st->dc_bwd[hi] + st->dr_fwd[lo] * st->dc_fwd[hi];
        sec = sec - st->dr_bwd[hi] * st->dc_bwd[lo] + st->dr_fwd[hi] * st->dc_fwd[lo];
        tt = (fir - mu[hi] * sumU_sumL[lo]) * invsig[hi];
        local_del = nt[lo] - 2 * tt;
        lb = 0.5 * (sqrt(2 * ((local_del + 2 * subseqlen * mu[lo] * mu[lo] * invsig[lo]) * invsig[lo] + 2 * subseqlen) -
                         DUL2[lo]) -
                    DUL[lo]);
        if (lb > bsf) {
            (*cnt)++;
            st->lb_vector[lo] = 1;
            continue;
        }
        tt = (sec - mu[lo] * sumU_sumL[hi]) * invsig[lo];
        local_del = nt[hi] - 2 * tt;
        lb = 0.5 * (sqrt(2 * ((local_del + 2 * subseqlen * mu[hi] * mu[hi] * invsig[hi]) * invsig[hi] + 2 * subseqlen) -
                         DUL2[hi]) -
                    DUL[hi]);
        if (lb > bsf) {
            st->lb_vector[lo] = 1;
            (*cnt)++;
        }
    }
}

/* the two LB_Kim end-point passes of CPU_version1/diag_mask_global.cpp:179-222 and
 * CPU_version1/lb_kim.cpp:28-71; lbn2 = lb_vector_new^2 (0 for LB_KIM) */
static int kim_prune(join_state *st, int64_t col, int64_t row, double threshold2, double lbn, int use_lbn,
                     double *cnt) {
    const double *t_ = SUB(st, row), *q = SUB(st, col);
    int m = st->m;
    double d;
    double x0 = t_[0];
    double y0 = t_[(m - 1)];
    const double dleft_orgin = ORC_DIST(x0, q[0]);
    const double dright_orgin = ORC_DIST(y0, q[m - 1]);
    double dleft = dleft_orgin, dright = dright_orgin;
    double x1 = (t_[(1)]);
    const double d_left_weak = smin(ORC_DIST(x1, q[0]), ORC_DIST(x1, q[1]));
    d = smin(d_left_weak, ORC_DIST(x0, q[1]));
    dleft += d;
    double y1 = (t_[(m - 2)]);
    const double d_right_weak = smin(ORC_DIST(y1, q[m - 1]), ORC_DIST(y1, q[m - 2]));
    d = smin(d_right_weak, ORC_DIST(y0, q[m - 2]));
    dright += d;
    if ((use_lbn ? (dleft + dright + lbn * lbn) : (dleft + dright)) >= threshold2) {
        (*cnt)++;
        st->lb_vector[col] = 1;
        return 1;
    }
    d = ORC_MIN(ORC_DIST(x1, q[0]) + ORC_DIST(t_[2], q[0]), d_left_weak + ORC_DIST(t_[2], q[1]));
    d = ORC_MIN(d, ORC_DIST(q[1], t_[1]) + ORC_DIST(q[2], t_[2]));
    d = ORC_MIN(d, dleft + ORC_DIST(q[2], t_[1]));
    d = ORC_MIN(d, ORC_DIST(q[1], t_[0]) + ORC_DIST(q[2], t_[0]));
    dleft = d + dleft_orgin;
    d = ORC_MIN(ORC_DIST(t_[m - 2], q[m - 1]) + ORC_DIST(t_[m - 3], q[m - 1]), d_right_weak + ORC_DIST(t_[m - 3], q[m - 2]));
    d = ORC_MIN(d, ORC_DIST(q[m - 2], t_[m - 2]) + ORC_DIST(q[m - 3], t_[m - 3]));
    d = ORC_MIN(d, dright + ORC_DIST(q[m - 3], t_[m - 2]));
    d = ORC_MIN(d, ORC_DIST(q[m - 2], t_[m - 1]) + ORC_DIST(q[m - 3], t_[m - 1]));
    dright = d + dright_orgin;
    if ((use_lbn ? (dleft + dright + lbn * lbn) : (dleft + dright)) >= threshold2) {
        st->lb_vector[col] = 1;
        (*cnt)++;
        return 1;
    }
    st->special[2 * col] = dleft;
    st->special[2 * col + 1] = dright;
    return 0;
}

/* CPU_version1/lb_kim.cpp:10-73 */
static void lb_kim(join_state *st, double threshold, int64_t temp_1, int diag, double *cnt) {
    for (int64_t col = 0; col < temp_1; col++) {
        int64_t row = diag + col - 1;
        if (st->lb_vector[col]) continue;
        kim_prune(st, col, row, threshold * threshold, 0.0, 0, cnt);
    }
}

/* CPU_version1/diag_mask_global.cpp:11-226 -- masked ball bound + LB_Kim end points */
static void diag_mask_global(join_state *st, int diag, double bsf, double *cnt, double *out_fir, double *out_sec) {
    const double *ts = st->a, *mu = st->mu, *invsig = st->invsig;
    const double *sumMASK = st->sumMASK_g, *sumU_sumL = st->sumU_sumL_g, *nUL = st->normUL_g;
    const double *DUL2 = st->DUL2_g, *DUL = st->DUL_g;
    int subseqlen = st->m;
    int64_t subcount = st->subcount;
    double M_NORM, t, dist;
    double cf = 0, d1f = 0, d2f = 0, cs = 0, d1s = 0, d2s = 0;
    for (int k = 0; k < subseqlen - 6; k++) cf += (ts + diag - 1 + 3)[k] * ((st->UTS_g + 3)[k] + (st->LTS_g + 3)[k]);
    for (int k = 0; k < subseqlen - 6; k++) d1f += (ts + diag - 1 + 3)[k] * (st->MASK + 3)[k];
    for (int k = 0; k < subseqlen - 6; k++) d2f += (st->TS2 + diag - 1 + 3)[k] * (st->MASK + 3)[k];
    for (int k = 0; k < subseqlen - 6; k++)
        cs += (ts + 3)[k] * ((st->UTS_g + diag - 1 + 3)[k] + (st->LTS_g + diag - 1 + 3)[k]);
    for (int k = 0; k < subseqlen - 6; k++) d1s += (ts + 3)[k] * (st->MASK + diag - 1 + 3)[k];
    for (int k = 0; k < subseqlen - 6; k++) d2s += (st->TS2 + 3)[k] * (st->MASK + diag - 1 + 3)[k];
    int64_t row, col;
    double lb;
    if (!st->lb_vector[0]) {
        row = 0;
        col = diag - 1;
        M_NORM = d2f - 2 * mu[col] * d1f + sumMASK[row] * mu[col] * mu[col];
        M_NORM = M_NORM * invsig[col] * invsig[col];
        t = (cf - mu[col] * sumU_sumL[row] + 2 * mu[row] * (sumMASK[row] * mu[col] - d1f)) * (invsig[col] * invsig[row]);
        dist = 2 * M_NORM - 2 * t + nUL[row];
        if (dist < DUL2[row])
            lb = 0;
        else
            lb = 0.5 * (sqrt(2 * dist - DUL2[row]) - DUL[row]);
        st->lb_vector_new[row] = lb;
        if (out_fir) out_fir[0] = lb;
        if (lb > bsf) {
            st->lb_vector[row] = 1;
            (*cnt)++;
        } else {
            col = 0;
            row = diag - 1;
            M_NORM = d2s - 2 * mu[col] * d1s + sumMASK[row] * mu[col] * mu[col];
            M_NORM = M_NORM * invsig[col] * invsig[col];
            t = (cs - mu[col] * sumU_sumL[row] + 2 * mu[row] * (sumMASK[row] * mu[col] - d1s)) *
                (invsig[col] * invsig[row]);
            /* the reference indexes the envelope norm with col here (:79); kept as is */
            dist = 2 * M_NORM - 2 * t + nUL[col];
            if (dist < DUL2[row])
                lb = 0;
            else
                lb = 0.5 * (sqrt(2 * dist - DUL2[row]) - DUL[row]);
            st->lb_vector_new[col] = lb;
            if (out_sec) out_sec[0] = lb;
            if (lb > bsf) {
                st->lb_vector[col] = 1;
                (*cnt)++;
            }
        }
    }
    for (int64_t lo = 1; lo < subcount - diag + 1; lo++) {
        int64_t hi = diag + lo - 1;
        cf = cf - st->dr_bwd_g[lo] * st->dc_bwd_g[hi] + (st->dr_fwd_g[lo]) * st->dc_fwd_g[hi];
        d1f = d1f - st->dr_bwdM[lo] * st->dc_bwd_g[hi] + st->dr_fwdM[lo] * st->dc_fwd_g[hi];
        d2f = d2f - st->dr_bwdM[lo] * st->dc_bwdT2[hi] + st->dr_fwdM[lo] * st->dc_fwdT2[hi];
        cs = cs - st->dr_bwd_g[hi] * st->dc_bwd_g[lo] + (st->dr_fwd_g[hi]) * st->dc_fwd_g[lo];
        d1s = d1s - st->dr_bwdM[hi] * st->dc_bwd_g[lo] + st->dr_fwdM[hi] * st->dc_fwd_g[lo];
        d2s = d2s - st->dr_bwdM[hi] * st->dc_bwdT2[lo] + st->dr_fwdM[hi] * st->dc_fwdT2[lo];
        if (st->lb_vector[lo]) continue;
        M_NORM = d2f - 2 * mu[hi] * d1f + sumMASK[lo] * mu[hi] * mu[hi];
        M_NORM = M_NORM * invsig[hi] * invsig[hi];
        t = (cf - mu[hi] * sumU_sumL[lo] + 2 * mu[lo] * (sumMASK[lo] * mu[hi] - d1f)) * (invsig[hi] * invsig[lo]);
        dist = 2 * M_NORM - 2 * t + nUL[lo];
        if (dist < DUL2[lo]) {
            lb = 0;
            if (out_fir) out_fir[lo] = lb;
        } else {
            lb = 0.5 * (sqrt(2 * dist - DUL2[lo]) - DUL[lo]);
            if (out_fir) out_fir[lo] = lb;
            st->lb_vector_new[lo] = lb;
            if (lb > bsf) {
                st->lb_vector[lo] = 1;
                (*cnt)++;
                continue;
            }
        }
        M_NORM = d2s - 2 * mu[lo] * d1s + sumMASK[hi] * mu[lo] * mu[lo];
        M_NORM = M_NORM * invsig[lo] * invsig[lo];
        t = (cs - mu[lo] * sumU_sumL[hi] + 2 * mu[hi] * (sumMASK[hi] * mu[lo] - d1s)) * (invsig[lo] * invsig[hi]);
        dist = 2 * M_NORM - 2 * t + nUL[hi];
        if (dist < DUL2[hi]) {
            lb = 0;
            if (out_sec) out_sec[lo] = lb;
        } else {
            lb = 0.5 * (sqrt(2 * dist - DUL2[hi]) - DUL[hi]);
            if (out_sec) out_sec[lo] = lb;
            st->lb_vector_new[lo] = smax(lb, st->lb_vector_new[lo]);
            if (lb > bsf) {
                st->lb_vector[lo] = 1;
                (*cnt)++;
            }
        }
    }
    int64_t temp_1 = subcount - diag + 1;
    double threshold2 = bsf * bsf;
    for (col = 0; col < temp_1; col++) {
        row = diag + col - 1;
        if (st->lb_vector[col]) continue;
        kim_prune(st, col, row, threshold2, st->lb_vector_new[col], 1, cnt);
    }
}

/* CPU_version1/compute_shared_data_global.cpp:12-46 and CPU_version1/compute_MASK_global.cpp:11-93 */
static void mask_global(join_state *st, int len_of_table, double *diff_out, double *ths_out) {
    int m = st->m;
    int64_t subcount = st->subcount, len = st->len;
    double *ct = (double *)calloc((size_t)m * len_of_table, sizeof(double));
    for (int64_t j = 0; j < subcount; j++) {
        for (int i = 0; i < m; i++) {
            double value = SUB(st, j)[i];
            int pp = (int)floor(value * 100 + len_of_table * 0.5);
            if (pp <= 0) pp = 1;
            if (pp > len_of_table) pp = len_of_table;
            ct[(size_t)i * len_of_table + pp - 1] = ct[(size_t)i * len_of_table + pp - 1] + 1;
        }
    }
    for (int i = 0; i < m; i++) {
        double sum1 = 0;
        for (int j = 0; j < len_of_table; j++) {
            sum1 += ct[(size_t)i * len_of_table + j];
            ct[(size_t)i * len_of_table + j] = sum1;
        }
    }
    double *nd = (double *)calloc((size_t)len, sizeof(double));
    for (int64_t i = 0; i < subcount; i++) {
        for (int j = 0; j < m; j++) {
            double LL = SUBL(st, i)[j] - 0.0, UU = SUBU(st, i)[j] + 0.0;
            double posLL = floor(LL * 100 + len_of_table * 0.5);
            double posUU = floor(UU * 100 + len_of_table * 0.5);
            if (posLL <= 0) posLL = 1;
            if (posLL > len_of_table) posLL = len_of_table;
            if (posUU <= 0) posUU = 1;
            if (posUU > len_of_table) posUU = len_of_table;
            double diff2 = ct[(size_t)j * len_of_table + (int64_t)posUU - 1] - ct[(size_t)j * len_of_table + (int64_t)posLL - 1];
            nd[i + j] += diff2 / subcount;
        }
    }
    /* divide by the coverage ramp mod = [1..m, m x (n-2m), m..1] */
    for (int64_t t = 0; t < len; t++) {
        double mod;
        if (t < m)
            mod = (double)(t + 1);
        else if (t < len - m)
            mod = (double)m;
        else
            mod = (double)m + (double)(t - (len - m)) * ((1.0 - m) / (m - 1)); /* linspace(m, 1, m) */
        nd[t] = nd[t] / mod;
    }
    double avg = 0.0;
    for (int64_t t = 0; t < len; t++) avg += nd[t];
    avg = avg / (double)len;
    double mean = 0.0;
    for (int64_t t = 0; t < len; t++) mean += nd[t];
    mean /= len;
    double ssd = 0.0;
    for (int64_t t = 0; t < len; t++) ssd += (nd[t] - mean) * (nd[t] - mean);
    double sd = sqrt(ssd / len);
    int adjust_factor = 0;
    double ths = avg + adjust_factor * sd;
    for (int64_t t = 0; t < len; t++) st->MASK[t] = (nd[t] <= ths) ? 1 : 0;
    if (diff_out) memcpy(diff_out, nd, sizeof(double) * (size_t)len);
    if (ths_out) *ths_out = ths;
    free(ct);
    free(nd);
}

static void *xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

/* CPU_version1/new_dtw_motifGUI.cpp:151-297: the global occupancy mask (table length from the
 * largest |z-value|) and the masked tables of diag_mask_global; needs the windows (subs) and
 * the raw-series envelope (UTS, LTS).  Optional outputs: the occupancy statistic per position,
 * the threshold and the table length. */
static void global_tables(join_state *st, double *diff_out, double *ths_out, int *len_out) {
    const double *a = st->a;
    const int m = st->m, subseqLen = st->m;
    const int64_t len = st->len, subcount = st->subcount;
    double *tmpv = (double *)xcalloc(len, sizeof(double));
    double *nL = (double *)xcalloc(subseqLen, sizeof(double)), *nU = (double *)xcalloc(subseqLen, sizeof(double));
    double real_max = SUB(st, 0)[0], real_min = SUB(st, 0)[0];
    for (int64_t r = 0; r < subcount; r++)
        for (int q = 0; q < m; q++) {
            double v2 = SUB(st, r)[q];
            if (v2 > real_max) real_max = v2;
            if (v2 < real_min) real_min = v2;
        }
    double max_real_value = floor(ORC_MAX(fabs(real_max), fabs(real_min))) * 2 * 100;
    int len_of_table_global = (int)(ORC_MAX(600, max_real_value));
    st->MASK = (double *)xcalloc(len, sizeof(double));
    mask_global(st, len_of_table_global, diff_out, ths_out);
    if (len_out) *len_out = len_of_table_global;
    st->TS2 = (double *)xcalloc(len, sizeof(double));
    for (int64_t x = 0; x < len; x++) st->TS2[x] = a[x] * a[x];
    st->UTS_g = (double *)xcalloc(len, sizeof(double));
    st->LTS_g = (double *)xcalloc(len, sizeof(double));
    for (int64_t x = 0; x < len; x++) {
        st->UTS_g[x] = st->UTS[x] * st->MASK[x];
        st->LTS_g[x] = st->LTS[x] * st->MASK[x];
    }
    /* shifted tables (BIAS = 3): index x >= 1 reads position x + 2 (backward) and
       x + m - 4 (forward) */
    st->dr_bwd_g = (double *)xcalloc(len, sizeof(double));
    st->dc_bwd_g = (double *)xcalloc(len, sizeof(double));
    st->dr_fwd_g = (double *)xcalloc(len, sizeof(double));
    st->dc_fwd_g = (double *)xcalloc(len, sizeof(double));
    st->dr_bwdM = (double *)xcalloc(len, sizeof(double));
    st->dr_fwdM = (double *)xcalloc(len, sizeof(double));
    st->dc_bwdT2 = (double *)xcalloc(len, sizeof(double));
    st->dc_fwdT2 = (double *)xcalloc(len, sizeof(double));
    st->dr_bwd_g[0] = 0.0 + 0.0;
    for (int64_t x = 1; x + 2 < len; x++) {
        st->dr_bwd_g[x] = st->UTS_g[x + 2] + st->LTS_g[x + 2];
        st->dc_bwd_g[x] = a[x + 2];
        st->dr_bwdM[x] = st->MASK[x + 2];
        st->dc_bwdT2[x] = st->TS2[x + 2];
    }
    for (int64_t x = 0; x + m - 4 < len; x++) {
        st->dr_fwd_g[x] = st->UTS_g[x + m - 4] + st->LTS_g[x + m - 4];
        st->dc_fwd_g[x] = a[x + m - 4];
        st->dr_fwdM[x] = st->MASK[x + m - 4];
        st->dc_fwdT2[x] = st->TS2[x + m - 4];
    }
    int b6 = subseqLen - 1 - 6;
    double *sUg = (double *)xcalloc(subcount, sizeof(double)), *sLg = (double *)xcalloc(subcount, sizeof(double));
    double *sU2g = (double *)xcalloc(subcount, sizeof(double)), *sL2g = (double *)xcalloc(subcount, sizeof(double));
    st->sumMASK_g = (double *)xcalloc(subcount, sizeof(double));
    movsum(st->UTS_g + 3, len - 6, b6, sUg);
    movsum(st->LTS_g + 3, len - 6, b6, sLg);
    for (int64_t x = 0; x < len - 6; x++) tmpv[x] = (st->UTS_g + 3)[x] * (st->UTS_g + 3)[x];
    movsum(tmpv, len - 6, b6, sU2g);
    for (int64_t x = 0; x < len - 6; x++) tmpv[x] = (st->LTS_g + 3)[x] * (st->LTS_g + 3)[x];
    movsum(tmpv, len - 6, b6, sL2g);
    movsum(st->MASK + 3, len - 6, b6, st->sumMASK_g);
    st->sumU_sumL_g = (double *)xcalloc(subcount, sizeof(double));
    double *sU2sL2g = (double *)xcalloc(subcount, sizeof(double));
    for (int64_t x = 0; x < subcount; x++) {
        sU2sL2g[x] = sU2g[x] + sL2g[x];
        st->sumU_sumL_g[x] = sUg[x] + sLg[x];
    }
    st->normUL_g = (double *)xcalloc(subcount, sizeof(double));
    for (int64_t row = 1; row <= subcount; row++) {
        double mu_ = st->mu[row - 1];
        st->normUL_g[row - 1] = (sU2sL2g[row - 1] - 2 * st->sumU_sumL_g[row - 1] * mu_ +
                                 2 * st->sumMASK_g[row - 1] * mu_ * mu_) *
                                st->invsig[row - 1] * st->invsig[row - 1];
    }
    st->DUL2_g = (double *)xcalloc(subcount, sizeof(double));
    st->DUL_g = (double *)xcalloc(subcount, sizeof(double));
    for (int64_t row = 0; row < subcount; row++) {
        for (int64_t i = row; i <= row + subseqLen - 1; i++) {
            nL[i - row] = (st->LTS_g[i] - st->mu[row]) * st->invsig[row];
            nU[i - row] = (st->UTS_g[i] - st->mu[row]) * st->invsig[row];
        }
        for (int q = 0; q < subseqLen; q++) tmpv[q] = nL[q] - nU[q];
        st->DUL2_g[row] = pow(norm2(tmpv, subseqLen), 2);
        st->DUL_g[row] = sqrt(ORC_MAX(st->DUL2_g[row], 0));
    }
    free(tmpv); free(nL); free(nU);
    free(sUg); free(sLg); free(sU2g); free(sL2g); free(sU2sL2g);
}

/* ---------------- the alpha-tuned local masks (CPU_version1/new_dtw_motifGUI.cpp:150-176,
 * 311-498, CPU_version1/compute_shared_data.cpp:10-135, CPU_version1/diag_mask_local_up.cpp:12-239,
 * CPU_version1/diag_mask_local_down.cpp:12-227) ---------------- */

/* CPU_version1/compute_shared_data.cpp:10-28: bins of 0.01 within [-3, 3], of 0.1 outside; the
 * double is truncated to int on assignment */
static int find_pos_local(double input, int length) {
    int pos;
    if (input > 3)
        pos = (int)(300 + (input - 3) * 10);
    else if (input < -3)
        pos = (int)(-300 + (input + 3) * 10);
    else
        pos = (int)(input * 100);
    pos += length / 2;
    return pos;
}

typedef struct {
    int len_of_cdf;
    short *ct;               /* len x len_of_cdf: cumulative counts of the z-values at each position */
    double *pos_UU, *pos_LL; /* len: bins of the mean envelope at each position */
    double *invsig_2;        /* subcount */
    /* vectors the reference allocates once per join (CPU_version1/new_dtw_motifGUI.cpp:311-326) and
       every local-mask call overwrites in part: rows a call does not write keep what an earlier call
       left there (zeros if none did) */
    double *sumU_sumL, *sumU2_sumL2, *sumMASK, *normUL;     /* subcount */
    double *raw_DIFF_UL, *raw_DIFF_UL2, *raw_DIFF_UL2_temp; /* len */
    double *DUL2_raw, *DUL2;                                /* subcount */
    double bad_bin;          /* reads of a cumulative count at bin 0 (index -1 in the reference) */
} local_state;

/* CPU_version1/new_dtw_motifGUI.cpp:150-176 with CPU_version1/compute_shared_data.cpp:30-135 */
static void local_tables(join_state *st, local_state *ls) {
    const int m = st->m;
    const int64_t len = st->len, subcount = st->subcount;
    memset(ls, 0, sizeof(*ls));
    double real_max = SUB(st, 0)[0], real_min = SUB(st, 0)[0];
    for (int64_t r = 0; r < subcount; r++)
        for (int q = 0; q < m; q++) {
            double v2 = SUB(st, r)[q];
            if (v2 > real_max) real_max = v2;
            if (v2 < real_min) real_min = v2;
        }
    double MAX_REAL_VALUE = ORC_MAX(fabs(real_max), fabs(real_min));
    int L;
    if (MAX_REAL_VALUE > 3)
        L = (int)((MAX_REAL_VALUE - 3) * 20 + 600);
    else
        L = 600;
    ls->len_of_cdf = L;
    short *ct = (short *)calloc((size_t)len * L, sizeof(short));
    ls->ct = ct;
    for (int64_t j = 0; j < subcount; j++)
        for (int i = 0; i < m; i++) {
            double value = SUB(st, j)[i];
            int pp = find_pos_local(value, L);
            if (pp < 0) pp = 0;
            if (pp >= L) pp = L - 1;
            ct[(size_t)(i + j) * L + pp] = (short)(ct[(size_t)(i + j) * L + pp] + 1);
        }
    for (int64_t i = 0; i < len; i++) {
        int sum1 = 0;
        for (int j = 0; j < L; j++) {
            sum1 = sum1 + ct[(size_t)i * L + j];
            ct[(size_t)i * L + j] = (short)sum1;
        }
    }
    double *sum_UU = (double *)calloc(len, sizeof(double)), *sum_LL = (double *)calloc(len, sizeof(double));
    double *cnt = (double *)calloc(len, sizeof(double));
    for (int64_t i = 0; i < subcount; i++)
        for (int j = 0; j < m; j++) {
            sum_UU[i + j] = sum_UU[i + j] + SUBU(st, i)[j];
            sum_LL[i + j] = sum_LL[i + j] + SUBL(st, i)[j];
            cnt[i + j] = cnt[i + j] + 1;
        }
    for (int64_t i = 0; i < len; i++) {
        sum_UU[i] = sum_UU[i] / cnt[i];
        sum_LL[i] = sum_LL[i] / cnt[i];
    }
    ls->pos_UU = (double *)calloc(len, sizeof(double));
    ls->pos_LL = (double *)calloc(len, sizeof(double));
    for (int64_t i = 0; i < len; i++) {
        double posLL = find_pos_local(sum_LL[i], L), posUU = find_pos_local(sum_UU[i], L);
        /* :113-135, including the check of posUU that sets posLL */
        if (posLL < 0) posLL = 1;
        if (posLL > L) posLL = L;
        if (posUU < 0) posLL = 1;
        if (posLL > L) posLL = L;
        if (posUU < 0) posUU = 1;
        if (posUU > L) posUU = L;
        ls->pos_UU[i] = posUU;
        ls->pos_LL[i] = posLL;
    }
    free(sum_UU);
    free(sum_LL);
    free(cnt);
    ls->invsig_2 = (double *)xcalloc(subcount, sizeof(double));
    for (int64_t i = 0; i < subcount; i++) ls->invsig_2[i] = st->invsig[i] * st->invsig[i];
    ls->sumU_sumL = (double *)xcalloc(subcount, sizeof(double));
    ls->sumU2_sumL2 = (double *)xcalloc(subcount, sizeof(double));
    ls->sumMASK = (double *)xcalloc(subcount, sizeof(double));
    ls->normUL = (double *)xcalloc(subcount, sizeof(double));
    ls->raw_DIFF_UL = (double *)xcalloc(len, sizeof(double));
    ls->raw_DIFF_UL2 = (double *)xcalloc(len, sizeof(double));
    ls->raw_DIFF_UL2_temp = (double *)xcalloc(len, sizeof(double));
    ls->DUL2_raw = (double *)xcalloc(subcount, sizeof(double));
    ls->DUL2 = (double *)xcalloc(subcount, sizeof(double));
}

static void local_free(local_state *ls) {
    free(ls->ct); free(ls->pos_UU); free(ls->pos_LL); free(ls->invsig_2); free(ls->sumU_sumL);
    free(ls->sumU2_sumL2); free(ls->sumMASK); free(ls->normUL); free(ls->raw_DIFF_UL); free(ls->raw_DIFF_UL2);
    free(ls->raw_DIFF_UL2_temp); free(ls->DUL2_raw); free(ls->DUL2);
}

/* count_table[pos][p - 1] - count_table[pos][q - 1]; a bin of 0 (the reference then reads index
 * -1, outside the row) is counted in bad_bin and read as a count of 0 */
static double local_count_diff(local_state *ls, int64_t pos, double p, double q) {
    const int L = ls->len_of_cdf;
    size_t pu = (size_t)p, pl = (size_t)q;
    int cu = 0, cl = 0;
    if (pu == 0) ls->bad_bin++; else cu = ls->ct[(size_t)pos * L + pu - 1];
    if (pl == 0) ls->bad_bin++; else cl = ls->ct[(size_t)pos * L + pl - 1];
    return (double)(cu - cl);
}

/* CPU_version1/diag_mask_local_up.cpp:12-239: the masked ball bound of the envelope of window
 * lo against window lo + diag - 1, the mask from the occupancy of each position's mean envelope
 * at the partner position, threshold (mean + 0.5) / 2 + alpha */
static void local_up(join_state *st, local_state *ls, double bsf, int diag, unsigned char *lb_vector,
                     double *cnt, double alpha, int64_t diag_bias) {
    const int subseqlen = st->m;
    const int64_t len = st->len, subcount = st->subcount;
    const double *ts = st->a, *mu = st->mu, *invsig = st->invsig, *TS2 = st->TS2;
    const double *UTS_ = st->UTS, *LTS_ = st->LTS;
    const int64_t nlen = len - diag_bias;
    double *normal_DIFF = (double *)xcalloc(nlen, sizeof(double));
    for (int64_t i = 0; i < nlen; i++) {
        int64_t pos = i + diag - 1;
        if (pos >= len) continue;
        double diff2 = local_count_diff(ls, pos, ls->pos_UU[i], ls->pos_LL[i]);
        double diff_ratio = diff2 / subseqlen;
        normal_DIFF[i] = diff_ratio;
    }
    double sum = 0.0;
    int count = 0;
    for (int64_t i = 0; i < nlen; i++) {
        if (!isnan(normal_DIFF[i])) {
            sum += normal_DIFF[i];
            count++;
        }
    }
    double avg_diff = (count == 0) ? NAN : sum / count;
    double ths_diff = (avg_diff + 0.5) / 2 + alpha;
    for (int64_t i = 0; i < nlen; i++) ls->raw_DIFF_UL[i] = UTS_[i] - LTS_[i];
    for (int64_t i = 0; i < nlen; i++) ls->raw_DIFF_UL2_temp[i] = ls->raw_DIFF_UL[i] * ls->raw_DIFF_UL[i];
    double *MASK = (double *)xcalloc(nlen, sizeof(double));
    for (int64_t i = 0; i < nlen; i++) MASK[i] = (normal_DIFF[i] <= ths_diff) ? 1 : 0;
    double *UTS = (double *)xcalloc(nlen, sizeof(double)), *LTS = (double *)xcalloc(nlen, sizeof(double));
    for (int64_t i = 0; i < nlen; i++) {
        UTS[i] = UTS_[i] * MASK[i];
        LTS[i] = LTS_[i] * MASK[i];
    }
    /* dr_bwd: (U + L)[3 + i]; dr_fwd: (U + L)[m - 4 + i] (the reference's arrays reach two entries
       past UTS at their ends, entries no pair reads) */
    const int64_t nbwd = len - 1 - diag_bias, nfwd = len - subseqlen + 1 - diag_bias;
    double *dr_bwd_less = (double *)xcalloc(nbwd, sizeof(double));
    double *dr_fwd = (double *)xcalloc(nfwd, sizeof(double));
    for (int64_t i = 0; i < nbwd && i + 3 < nlen; i++) dr_bwd_less[i] = UTS[3 + i] + LTS[3 + i];
    for (int64_t i = 0; i < nfwd; i++) dr_fwd[i] = UTS[subseqlen - 4 + i] + LTS[subseqlen - 4 + i];
    double sum_sumU = 0, sum_sumL = 0, sum_sumU2 = 0, sum_sumL2 = 0, sum_sumMASK = 0;
    const double *U_pos = UTS + 3, *L_pos = LTS + 3, *MASK_pos = MASK + 3;
    double fU, fL, sU, sL;
    for (int i = 0; i < subseqlen - 6; i++) {
        fU = U_pos[i];
        fL = L_pos[i];
        sum_sumU += fU;
        sum_sumL += fL;
        sum_sumU2 += fU * fU;
        sum_sumL2 += fL * fL;
        sum_sumMASK += MASK_pos[i];
    }
    for (int64_t i = 0; i < subcount - 6 - diag_bias; i++) {
        ls->sumU_sumL[i] = sum_sumU + sum_sumL;
        ls->sumU2_sumL2[i] = sum_sumU2 + sum_sumL2;
        ls->sumMASK[i] = sum_sumMASK;
        fU = U_pos[i];
        fL = L_pos[i];
        sU = U_pos[i + subseqlen - 6];
        sL = L_pos[i + subseqlen - 6];
        sum_sumU = sum_sumU - fU + sU;
        sum_sumL = sum_sumL - fL + sL;
        sum_sumU2 = sum_sumU2 - fU * fU + sU * sU;
        sum_sumL2 = sum_sumL2 - fL * fL + sL * sL;
        sum_sumMASK = sum_sumMASK - MASK_pos[i] + MASK_pos[i + subseqlen - 6];
    }
    for (int64_t row = 0; row < subcount - diag_bias; row++)
        ls->normUL[row] = (ls->sumU2_sumL2[row] - 2 * ls->sumU_sumL[row] * mu[row] +
                           2 * ls->sumMASK[row] * mu[row] * mu[row]) * invsig[row] * invsig[row];
    for (int64_t i = 0; i < nlen; i++) ls->raw_DIFF_UL2[i] = ls->raw_DIFF_UL2_temp[i] * MASK[i];
    movsum(ls->raw_DIFF_UL2 + 3, nlen - 6, subseqlen - 7, ls->DUL2_raw);
    for (int64_t i = 0; i < ORC_MIN(subcount, nlen); i++) ls->DUL2[i] = ls->DUL2_raw[i] * ls->invsig_2[i];
    const double *dc_bwd_less = ts + 3, *dc_fwd = ts + subseqlen - 4;
    const double *dr_bwdMASK_less = MASK + 3, *dr_fwdMASK = MASK + subseqlen - 4;
    const double *dc_bwdTS2_less = TS2 + 3, *dc_fwdTS2 = TS2 + subseqlen - 4;
    double cov = 0, dTS = 0, dTS2 = 0;
    for (int k = 0; k < subseqlen - 6; k++) cov += (ts + diag - 1 + 3)[k] * ((UTS + 3)[k] + (LTS + 3)[k]);
    for (int k = 0; k < subseqlen - 6; k++) dTS += (ts + diag - 1 + 3)[k] * (MASK + 3)[k];
    for (int k = 0; k < subseqlen - 6; k++) dTS2 += (TS2 + diag - 1 + 3)[k] * (MASK + 3)[k];
    const double *sumMASK = ls->sumMASK, *sumU_sumL = ls->sumU_sumL, *nUL = ls->normUL, *DUL2 = ls->DUL2;
    double M_NORM, t, dist, lb;
    if (!lb_vector[0]) {
        int64_t row = 0, col = diag - 1;
        M_NORM = dTS2 - 2 * mu[col] * dTS + sumMASK[row] * mu[col] * mu[col];
        M_NORM = M_NORM * invsig[col] * invsig[col];
        t = (cov - mu[col] * sumU_sumL[row] + 2 * mu[row] * (sumMASK[row] * mu[col] - dTS)) * (invsig[col] * invsig[row]);
        dist = 2 * M_NORM - 2 * t + nUL[row];
        if (dist < DUL2[row])
            lb = 0;
        else
            lb = 0.5 * (sqrt(2 * dist - DUL2[row]) - sqrt(ORC_MAX(DUL2[row], 0)));
        if (lb > bsf || (lb * lb + st->special[2 * row] + st->special[2 * row + 1] > bsf * bsf)) {
            lb_vector[row] = 1;
            (*cnt)++;
        }
    }
    double bsf_2 = bsf * bsf;
    int64_t temp_1 = subcount - diag + 1;
    for (int64_t lo = 1; lo < temp_1; lo++) {
        int64_t hi = diag + lo - 1, lo_l = lo - 1, hi_l = hi - 1;
        cov = cov - dr_bwd_less[lo_l] * dc_bwd_less[hi_l] + dr_fwd[lo] * dc_fwd[hi];
        dTS = dTS - dr_bwdMASK_less[lo_l] * dc_bwd_less[hi_l] + dr_fwdMASK[lo] * dc_fwd[hi];
        dTS2 = dTS2 - dr_bwdMASK_less[lo_l] * dc_bwdTS2_less[hi_l] + dr_fwdMASK[lo] * dc_fwdTS2[hi];
        if (lb_vector[lo]) continue;
        M_NORM = dTS2 - 2 * mu[hi] * dTS + sumMASK[lo] * mu[hi] * mu[hi];
        M_NORM = M_NORM * invsig[hi] * invsig[hi];
        t = (cov - mu[hi] * sumU_sumL[lo] + 2 * mu[lo] * (sumMASK[lo] * mu[hi] - dTS)) * (invsig[hi] * invsig[lo]);
        dist = 2 * M_NORM - 2 * t + nUL[lo];
        if (dist < DUL2[lo]) {
            lb = 0;
        } else {
            lb = 0.5 * (sqrt(2 * dist - DUL2[lo]) - sqrt(ORC_MAX(DUL2[lo], 0)));
            if (lb > bsf || (lb * lb + st->special[2 * lo] + st->special[2 * lo + 1] > bsf_2)) {
                (*cnt)++;
                lb_vector[lo] = 1;
            }
        }
    }
    free(normal_DIFF); free(MASK); free(UTS); free(LTS); free(dr_bwd_less); free(dr_fwd);
}

/* CPU_version1/diag_mask_local_down.cpp:12-227: the same with the envelope of window lo + diag - 1
 * against window lo, the mask from positions from diag_bias on (positions below count as 0 in
 * its mean) */
static void local_down(join_state *st, local_state *ls, double bsf, int diag, unsigned char *lb_vector,
                       double *cnt, double alpha, int64_t diag_bias) {
    const int subseqlen = st->m;
    const int64_t len = st->len, subcount = st->subcount;
    const double *ts = st->a, *mu = st->mu, *invsig = st->invsig, *TS2 = st->TS2;
    const double *UTS_ = st->UTS, *LTS_ = st->LTS;
    double *normal_DIFF = (double *)xcalloc(len, sizeof(double));
    for (int64_t i = diag_bias; i < len; i++) {
        int64_t pos = i - diag + 1;
        if (pos < 0) continue;
        double diff2 = local_count_diff(ls, pos, ls->pos_UU[i], ls->pos_LL[i]);
        double diff_ratio = diff2 / subseqlen;
        normal_DIFF[i] = diff_ratio;
    }
    double sum = 0.0;
    int count = 0;
    for (int64_t i = 0; i < len; i++) {
        if (!isnan(normal_DIFF[i])) {
            sum += normal_DIFF[i];
            count++;
        }
    }
    double avg_diff = (count == 0) ? NAN : sum / count;
    double ths_diff = (avg_diff + 0.5) / 2 + alpha;
    for (int64_t i = diag_bias; i < len; i++) ls->raw_DIFF_UL[i] = UTS_[i] - LTS_[i];
    for (int64_t i = diag_bias; i < len; i++) ls->raw_DIFF_UL2_temp[i] = ls->raw_DIFF_UL[i] * ls->raw_DIFF_UL[i];
    double *MASK = (double *)xcalloc(len, sizeof(double));
    for (int64_t i = diag_bias; i < len; i++) MASK[i] = (normal_DIFF[i] <= ths_diff) ? 1 : 0;
    double *UTS = (double *)xcalloc(len, sizeof(double)), *LTS = (double *)xcalloc(len, sizeof(double));
    for (int64_t i = diag_bias; i < len; i++) {
        UTS[i] = UTS_[i] * MASK[i];
        LTS[i] = LTS_[i] * MASK[i];
    }
    const int64_t nbwd = len - 1, nfwd = len - subseqlen + 1;
    double *dr_bwd_less = (double *)xcalloc(nbwd, sizeof(double));
    double *dr_fwd = (double *)xcalloc(nfwd, sizeof(double));
    for (int64_t i = diag_bias; i < nbwd && i + 3 < len; i++) dr_bwd_less[i] = UTS[3 + i] + LTS[3 + i];
    for (int64_t i = diag_bias; i < nfwd; i++) dr_fwd[i] = UTS[subseqlen - 4 + i] + LTS[subseqlen - 4 + i];
    double sum_sumU = 0, sum_sumL = 0, sum_sumU2 = 0, sum_sumL2 = 0, sum_sumMASK = 0;
    const double *U_pos = UTS + 3, *L_pos = LTS + 3, *MASK_pos = MASK + 3;
    double fU, fL, sU, sL;
    for (int64_t i = diag_bias; i < diag_bias + subseqlen - 6; i++) {
        fU = U_pos[i];
        fL = L_pos[i];
        sum_sumU += fU;
        sum_sumL += fL;
        sum_sumU2 += fU * fU;
        sum_sumL2 += fL * fL;
        sum_sumMASK += MASK_pos[i];
    }
    for (int64_t i = diag_bias; i < subcount - 6; i++) {
        ls->sumU_sumL[i] = sum_sumU + sum_sumL;
        ls->sumU2_sumL2[i] = sum_sumU2 + sum_sumL2;
        ls->sumMASK[i] = sum_sumMASK;
        fU = U_pos[i];
        fL = L_pos[i];
        sU = U_pos[i + subseqlen - 6];
        sL = L_pos[i + subseqlen - 6];
        sum_sumU = sum_sumU - fU + sU;
        sum_sumL = sum_sumL - fL + sL;
        sum_sumU2 = sum_sumU2 - fU * fU + sU * sU;
        sum_sumL2 = sum_sumL2 - fL * fL + sL * sL;
        sum_sumMASK = sum_sumMASK - MASK_pos[i] + MASK_pos[i + subseqlen - 6];
    }
    for (int64_t row = diag_bias; row < subcount; row++)
        ls->normUL[row] = (ls->sumU2_sumL2[row] - 2 * ls->sumU_sumL[row] * mu[row] +
                           2 * ls->sumMASK[row] * mu[row] * mu[row]) * invsig[row] * invsig[row];
    for (int64_t i = diag_bias; i < len; i++) ls->raw_DIFF_UL2[i] = ls->raw_DIFF_UL2_temp[i] * MASK[i];
    {
        /* movsum_p_from_pos (CPU_version1/allmatrixop/movsum.cpp): sums from position diag_bias */
        const double *x = ls->raw_DIFF_UL2 + 3;
        const int64_t b = subseqlen - 7, n2 = (len - 6) - b;
        double sm = 0;
        for (int64_t i = diag_bias; i <= b + diag_bias; i++) sm += x[i];
        for (int64_t i = diag_bias; i < n2 - 1; i++) {
            ls->DUL2_raw[i] = sm;
            sm = sm - x[i] + x[i + b + 1];
        }
        ls->DUL2_raw[n2 - 1] = sm;
    }
    for (int64_t i = diag_bias; i < subcount; i++) ls->DUL2[i] = ls->DUL2_raw[i] * ls->invsig_2[i];
    const double *dc_bwd_less = ts + 3, *dc_fwd = ts + subseqlen - 4;
    const double *dr_bwdMASK_less = MASK + 3, *dr_fwdMASK = MASK + subseqlen - 4;
    const double *dc_bwdTS2_less = TS2 + 3, *dc_fwdTS2 = TS2 + subseqlen - 4;
    double cov = 0, dTS = 0, dTS2 = 0;
    for (int k = 0; k < subseqlen - 6; k++) cov += (ts + 3)[k] * ((UTS + diag - 1 + 3)[k] + (LTS + diag - 1 + 3)[k]);
    for (int k = 0; k < subseqlen - 6; k++) dTS += (ts + 3)[k] * (MASK + diag - 1 + 3)[k];
    for (int k = 0; k < subseqlen - 6; k++) dTS2 += (TS2 + 3)[k] * (MASK + diag - 1 + 3)[k];
    const double *sumMASK = ls->sumMASK, *sumU_sumL = ls->sumU_sumL, *nUL = ls->normUL, *DUL2 = ls->DUL2;
    double M_NORM, t, dist, lb;
    if (!lb_vector[0]) {
        int64_t col = 0, row = diag - 1;
        M_NORM = dTS2 - 2 * mu[col] * dTS + sumMASK[row] * mu[col] * mu[col];
        M_NORM = M_NORM * invsig[col] * invsig[col];
        t = (cov - mu[col] * sumU_sumL[row] + 2 * mu[row] * (sumMASK[row] * mu[col] - dTS)) * (invsig[col] * invsig[row]);
        dist = 2 * M_NORM - 2 * t + nUL[row];
        if (dist < DUL2[row])
            lb = 0;
        else
            lb = 0.5 * (sqrt(2 * dist - DUL2[row]) - sqrt(ORC_MAX(DUL2[row], 0)));
        if (lb > bsf || (lb * lb + st->special[2 * col] + st->special[2 * col + 1] > bsf * bsf)) {
            lb_vector[col] = 1;
            (*cnt)++;
        }
    }
    double bsf_2 = bsf * bsf;
    int64_t temp_1 = subcount - diag + 1;
    for (int64_t lo = 1; lo < temp_1; lo++) {
        int64_t hi = diag + lo - 1, lo_l = lo - 1, hi_l = hi - 1;
        cov = cov - dr_bwd_less[hi_l] * dc_bwd_less[lo_l] + dr_fwd[hi] * dc_fwd[lo];
        dTS = dTS - dr_bwdMASK_less[hi_l] * dc_bwd_less[lo_l] + dr_fwdMASK[hi] * dc_fwd[lo];
        dTS2 = dTS2 - dr_bwdMASK_less[hi_l] * dc_bwdTS2_less[lo_l] + dr_fwdMASK[hi] * dc_fwdTS2[lo];
        if (lb_vector[lo]) continue;
        M_NORM = dTS2 - 2 * mu[lo] * dTS + sumMASK[hi] * mu[lo] * mu[lo];
        M_NORM = M_NORM * invsig[lo] * invsig[lo];
        t = (cov - mu[lo] * sumU_sumL[hi] + 2 * mu[hi] * (sumMASK[hi] * mu[lo] - dTS)) * (invsig[lo] * invsig[hi]);
        dist = 2 * M_NORM - 2 * t + nUL[hi];
        if (dist < DUL2[hi]) {
            lb = 0;
        } else {
            lb = 0.5 * (sqrt(2 * dist - DUL2[hi]) - sqrt(ORC_MAX(DUL2[hi], 0)));
            if (lb > bsf || (lb * lb + st->special[2 * lo] + st->special[2 * lo + 1] > bsf_2)) {
                (*cnt)++;
                lb_vector[lo] = 1;
            }
        }
    }
    free(normal_DIFF); free(MASK); free(UTS); free(LTS); free(dr_bwd_less); free(dr_fwd);
}

/* The state global_tables and diag_mask_global need, with the caller's window statistics (or
 * the reference's movmean when mu is NULL): z-normalised windows and their clipped envelopes
 * (CPU_version1/new_dtw_motifGUI.cpp:31-52) and the raw-series envelope (:96-98). */
static int mask_state(join_state *st, const double *a, int64_t n, int32_t m, int32_t w, const double *mu,
                      const double *sig) {
    memset(st, 0, sizeof(*st));
    if (m < 8 || w < 0 || n < 2 * (int64_t)m + 1) return 1;
    st->len = n;
    st->subcount = n - m + 1;
    st->m = m;
    st->w = w;
    st->a = a;
    st->mu = (double *)xcalloc(st->subcount, sizeof(double));
    st->sig = (double *)xcalloc(st->subcount, sizeof(double));
    st->invsig = (double *)xcalloc(st->subcount, sizeof(double));
    if (mu && sig) {
        memcpy(st->mu, mu, sizeof(double) * (size_t)st->subcount);
        memcpy(st->sig, sig, sizeof(double) * (size_t)st->subcount);
    } else {
        orc_mvmean(a, n, m, st->mu, st->sig);
    }
    for (int64_t i = 0; i < st->subcount; i++) {
        if (!(st->sig[i] > 0)) return 2;
        st->invsig[i] = 1 / st->sig[i];
    }
    st->subs = (double *)xcalloc((size_t)st->subcount * m, sizeof(double));
    st->subsU = (double *)xcalloc((size_t)st->subcount * m, sizeof(double));
    st->subsL = (double *)xcalloc((size_t)st->subcount * m, sizeof(double));
    for (int64_t i = 0; i < st->subcount; i++) {
        for (int j = 0; j < m; j++) SUB(st, i)[j] = (a[i + j] - st->mu[i]) / st->sig[i];
        orc_lemire(SUB(st, i), m, w, SUBL(st, i), SUBU(st, i));
    }
    st->UTS = (double *)xcalloc(n, sizeof(double));
    st->LTS = (double *)xcalloc(n, sizeof(double));
    orc_lemire(a, n, w, st->LTS, st->UTS);
    return 0;
}

static void mask_state_free(join_state *st) {
    free(st->mu); free(st->sig); free(st->invsig); free(st->subs); free(st->subsU); free(st->subsL);
    free(st->UTS); free(st->LTS); free(st->MASK); free(st->TS2); free(st->UTS_g); free(st->LTS_g);
    free(st->dr_bwd_g); free(st->dc_bwd_g); free(st->dr_fwd_g); free(st->dc_fwd_g); free(st->dr_bwdM);
    free(st->dr_fwdM); free(st->dc_bwdT2); free(st->dc_fwdT2); free(st->sumMASK_g); free(st->sumU_sumL_g);
    free(st->normUL_g); free(st->DUL2_g); free(st->DUL_g); free(st->special); free(st->lb_vector);
    free(st->lb_vector_new);
}

int orc_global_mask(const double *a, int64_t n, int32_t m, int32_t w, const double *mu, const double *sig,
                    unsigned char *mask, double *diff, double *ths, int32_t *len_of_table) {
    join_state S;
    int rc = mask_state(&S, a, n, m, w, mu, sig);
    if (rc == 0) {
        int len = 0;
        global_tables(&S, diff, ths, &len);
        for (int64_t t = 0; t < n; t++) mask[t] = S.MASK[t] != 0.0;
        if (len_of_table) *len_of_table = len;
    }
    mask_state_free(&S);
    return rc;
}

int orc_masked_ball_diag(const double *a, int64_t n, int32_t m, int32_t w, const double *mu, const double *sig,
                         int64_t diag, double *lb_fir, double *lb_sec) {
    join_state S;
    int rc = mask_state(&S, a, n, m, w, mu, sig);
    if (rc == 0 && (diag <= m || diag > S.subcount)) rc = 3;
    if (rc == 0) {
        global_tables(&S, NULL, NULL, NULL);
        S.special = (double *)xcalloc((size_t)S.subcount * 2, sizeof(double));
        S.lb_vector = (unsigned char *)xcalloc(S.subcount, 1);
        S.lb_vector_new = (double *)xcalloc(S.subcount, sizeof(double));
        double cnt = 0;
        /* with an infinite best-so-far nothing is pruned, so both directions run on every pair */
        diag_mask_global(&S, (int)diag, INFINITY, &cnt, lb_fir, lb_sec);
    }
    mask_state_free(&S);
    return rc;
}

/* CPU_version1/new_dtw_motifGUI.cpp:21-586 */
int orc_join(const double *a, int64_t n, int32_t m, int32_t w, const double *mp_ed, const int32_t *mp_ed_index,
             orc_join_result *out) {
    if (m < 8 || w < 0 || n < 2 * (int64_t)m + 1) return 1;
    memset(out, 0, sizeof(*out));
    join_state S;
    memset(&S, 0, sizeof(S));
    join_state *st = &S;
    int subseqLen = m, maxwarp = w;
    int64_t len = n, subcount = n - m + 1;
    st->len = len;
    st->subcount = subcount;
    st->m = m;
    st->w = w;
    st->a = a;
    st->mu = (double *)xcalloc(subcount, sizeof(double));
    st->sig = (double *)xcalloc(subcount, sizeof(double));
    st->invsig = (double *)xcalloc(subcount, sizeof(double));
    orc_mvmean(a, n, m, st->mu, st->sig);
    for (int64_t i = 0; i < subcount; i++) {
        if (!(st->sig[i] > 0)) { /* constant window: the reference divides by zero here */
            free(st->mu); free(st->sig); free(st->invsig);
            return 2;
        }
        st->invsig[i] = 1 / st->sig[i];
    }
    st->subs = (double *)xcalloc((size_t)subcount * m, sizeof(double));
    st->subsU = (double *)xcalloc((size_t)subcount * m, sizeof(double));
    st->subsL = (double *)xcalloc((size_t)subcount * m, sizeof(double));
    for (int64_t i = 0; i < subcount; i++) {
        for (int j = 0; j < m; j++) SUB(st, i)[j] = (a[i + j] - st->mu[i]) / st->sig[i];
        orc_lemire(SUB(st, i), m, w, SUBL(st, i), SUBU(st, i));
    }
    /* seed, :53-77 */
    int k = 100;
    int32_t *min_index = (int32_t *)xcalloc(k + 10, sizeof(int32_t));
    orc_min_v_k(mp_ed, subcount, k, min_index);
    double best_so_far = mp_ed[min_index[0]];
    out->mp_seed = best_so_far;
    double *cb_temp = (double *)xcalloc(subseqLen + maxwarp + 1, sizeof(double));
    int64_t low_index = 0, high_index = 0;
    int64_t seed_i = min_index[0], seed_j = (int64_t)mp_ed_index[min_index[0]] - 1;
    for (int i = 0; i < k; i++) {
        low_index = min_index[i];
        high_index = (int64_t)mp_ed_index[min_index[i]] - 1;
        if (lb_kk_first(SUB(st, low_index), SUB(st, high_index), SUBU(st, high_index), SUBL(st, low_index), subseqLen,
                        best_so_far * best_so_far, cb_temp)) {
            double dtw_value = sqrt(orc_mon_dtw(SUB(st, low_index), SUB(st, high_index), cb_temp, subseqLen, maxwarp,
                                                best_so_far * best_so_far));
            if (dtw_value < best_so_far) {
                best_so_far = dtw_value;
                seed_i = low_index;
                seed_j = high_index;
            }
        }
    }
    double first_min = ORC_MIN(low_index, high_index);
    double sec_min = ORC_MAX(low_index, high_index);
    double bsf = best_so_far;
    out->seed_bsf = best_so_far;
    /* raw envelope and diag_fast tables, :96-147 */
    st->UTS = (double *)xcalloc(len, sizeof(double));
    st->LTS = (double *)xcalloc(len, sizeof(double));
    orc_lemire(a, len, maxwarp, st->LTS, st->UTS);
    st->dr_bwd = (double *)xcalloc(subcount, sizeof(double));
    st->dc_bwd = (double *)xcalloc(subcount, sizeof(double));
    st->dr_fwd = (double *)xcalloc(subcount, sizeof(double));
    st->dc_fwd = (double *)xcalloc(subcount, sizeof(double));
    st->dr_bwd[0] = 0.0 + 0.0;
    st->dc_bwd[0] = 0.0;
    for (int64_t x = 1; x < subcount; x++) {
        st->dr_bwd[x] = st->UTS[x - 1] + st->LTS[x - 1];
        st->dc_bwd[x] = a[x - 1];
    }
    for (int64_t x = 0; x < subcount; x++) {
        st->dr_fwd[x] = st->UTS[x + m - 1] + st->LTS[x + m - 1];
        st->dc_fwd[x] = a[x + m - 1];
    }
    double *sumU = (double *)xcalloc(subcount, sizeof(double)), *sumL = (double *)xcalloc(subcount, sizeof(double));
    double *sumU2 = (double *)xcalloc(subcount, sizeof(double)), *sumL2 = (double *)xcalloc(subcount, sizeof(double));
    double *tmpv = (double *)xcalloc(len, sizeof(double));
    movsum(st->UTS, len, subseqLen - 1, sumU);
    movsum(st->LTS, len, subseqLen - 1, sumL);
    for (int64_t x = 0; x < len; x++) tmpv[x] = st->UTS[x] * st->UTS[x];
    movsum(tmpv, len, subseqLen - 1, sumU2);
    for (int64_t x = 0; x < len; x++) tmpv[x] = st->LTS[x] * st->LTS[x];
    movsum(tmpv, len, subseqLen - 1, sumL2);
    st->sumU_sumL = (double *)xcalloc(subcount, sizeof(double));
    double *sumU2_sumL2 = (double *)xcalloc(subcount, sizeof(double));
    for (int64_t x = 0; x < subcount; x++) {
        sumU2_sumL2[x] = sumU2[x] + sumL2[x];
        st->sumU_sumL[x] = sumU[x] + sumL[x];
    }
    st->DUL2_fast = (double *)xcalloc(subcount, sizeof(double));
    st->DUL_fast = (double *)xcalloc(subcount, sizeof(double));
    double *nL = (double *)xcalloc(subseqLen, sizeof(double)), *nU = (double *)xcalloc(subseqLen, sizeof(double));
    for (int64_t row = 0; row < subcount; row++) {
        for (int64_t i = row; i <= row + subseqLen - 1; i++) {
            nL[i - row] = (st->LTS[i] - st->mu[row]) * st->invsig[row];
            nU[i - row] = (st->UTS[i] - st->mu[row]) * st->invsig[row];
        }
        double s = 0.0;
        for (int q = 0; q < subseqLen; q++) s += nL[q] * nU[q];
        st->DUL2_fast[row] = pow(norm2(nL, subseqLen), 2) + pow(norm2(nU, subseqLen), 2) - 2 * s;
        st->DUL_fast[row] = pow(ORC_MAX(st->DUL2_fast[row], 0.0), 0.5);
    }
    st->normUL_trans = (double *)xcalloc(subcount, sizeof(double));
    for (int64_t row = 0; row < subcount; row++)
        st->normUL_trans[row] = (sumU2_sumL2[row] - 2 * st->sumU_sumL[row] * st->mu[row]) * st->invsig[row];
    /* global occupancy mask and masked tables, :151-297 */
    global_tables(st, NULL, NULL, NULL);
    st->special = (double *)xcalloc((size_t)subcount * 2, sizeof(double));
    st->lb_vector = (unsigned char *)xcalloc(subcount, 1);
    st->lb_vector_new = (double *)xcalloc(subcount, sizeof(double));
    /* the sweep, :300-558 */
    int minlag = subseqLen;
    double fast_prune_cnt = 0, global_prune_cnt = 0, kk_cnt = 0, dtw_cnt = 0;
    long long all_cnt = 0;
    int buffer_length = ORC_MAX(subseqLen / 10, 30);
    double *cb = (double *)xcalloc(subseqLen + maxwarp + 1, sizeof(double));
    double *cb1 = (double *)xcalloc(subseqLen + maxwarp, sizeof(double));
    double *cb2 = (double *)xcalloc(subseqLen + maxwarp, sizeof(double));
    double *proj = (double *)xcalloc(subseqLen, sizeof(double));
    double *proj_u = (double *)xcalloc(subseqLen, sizeof(double)), *proj_l = (double *)xcalloc(subseqLen, sizeof(double));
    double pe_cnt = 0, local_rounds = 0, local_down_rounds = 0, local_prune_cnt = 0;
    int found = 0;
    /* the alpha search of the local masks, :332-357 */
    local_state LS;
    local_tables(st, &LS);
    int alpha_interval = ORC_MAX(subseqLen * 5, 1000);
    int buffer_div_alpha = (int)ceil(1.0 * alpha_interval / buffer_length);
    static const double alpha_list[15] = {-0.7, -0.6, -0.5, -0.4, -0.3, -0.2, -0.1, 0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7};
    const double ALPHA_BIAS = 0.1; /* CPU_version1/alldef/matrix.h:16 */
    double *alpha_table = (double *)xcalloc(subcount, sizeof(double));
    unsigned char *lb_vector_for_next = (unsigned char *)xcalloc(subcount, 1);
    double best_alpha = 0;
    for (int i_buffer = 0; i_buffer < buffer_length; i_buffer++) {
        int i_table_num = 0;
        for (int64_t diag = minlag + 1 + i_buffer; diag <= subcount; diag += buffer_length) {
            i_table_num++;
            out->fast_stage_rounds++;
            memset(st->lb_vector_new, 0, sizeof(double) * subcount);
            memset(st->lb_vector, 0, subcount);
            double cnt_of_purn = 0;
            diag_fast(st, (int)diag, &cnt_of_purn, bsf);
            double cnt1 = cnt_of_purn;
            fast_prune_cnt += cnt1;
            double pruning_rate = cnt_of_purn / (double)(subcount - diag + 1);
            if (pruning_rate < 0.98) {
                out->global_rounds++;
                diag_mask_global(st, (int)diag, bsf, &cnt_of_purn, NULL, NULL);
            } else {
                lb_kim(st, bsf, subcount - diag + 1, (int)diag, &cnt_of_purn);
            }
            double cnt2 = cnt_of_purn;
            global_prune_cnt += (cnt2 - cnt1);
            pruning_rate = cnt_of_purn / (double)(subcount - diag + 1);
            /* :400-498: alpha of this diagonal (a search over alpha_list every buffer_div_alpha
               diagonals of the first pass, a local search or an extrapolation in between, the
               table interpolated on later passes), then the local masks, up and down, each while
               fewer than 80 % of the diagonal's pairs are pruned */
            {
                const int64_t diag_bias = diag - minlag - 1;
                double temp_pruning_cnt = 0;
                int best_pruning_cnt = 0;
                const double temp_alpha_list[5] = {best_alpha - 2 * ALPHA_BIAS, best_alpha - ALPHA_BIAS, best_alpha,
                                                   best_alpha + ALPHA_BIAS, best_alpha + 2 * ALPHA_BIAS};
                if (i_buffer == 0) {
                    if (i_table_num % buffer_div_alpha == 0) {
                        for (int q = 0; q < 15; q++) {
                            temp_pruning_cnt = 0;
                            memcpy(lb_vector_for_next, st->lb_vector, subcount);
                            local_up(st, &LS, bsf, (int)diag, lb_vector_for_next, &temp_pruning_cnt, alpha_list[q],
                                     diag_bias);
                            if (temp_pruning_cnt > best_pruning_cnt) {
                                best_pruning_cnt = (int)temp_pruning_cnt;
                                best_alpha = alpha_list[q];
                            }
                        }
                        alpha_table[i_table_num] = best_alpha;
                    } else {
                        if (pruning_rate < 0.85) {
                            for (int q = 0; q < 5; q++) {
                                temp_pruning_cnt = 0;
                                memcpy(lb_vector_for_next, st->lb_vector, subcount);
                                local_up(st, &LS, bsf, (int)diag, lb_vector_for_next, &temp_pruning_cnt,
                                         temp_alpha_list[q], diag_bias);
                                if (temp_pruning_cnt > best_pruning_cnt) {
                                    best_pruning_cnt = (int)temp_pruning_cnt;
                                    best_alpha = temp_alpha_list[q];
                                }
                            }
                            alpha_table[i_table_num] = best_alpha;
                        } else {
                            alpha_table[i_table_num] = 2 * alpha_table[ORC_MAX(i_table_num - 1, 0)] -
                                                       alpha_table[ORC_MAX(i_table_num - 2, 0)];
                        }
                    }
                }
                double alpha = alpha_table[i_table_num] +
                               (alpha_table[i_table_num + 1] - alpha_table[i_table_num]) / buffer_length * (i_buffer);
                if (pruning_rate < 0.80) {
                    local_rounds++;
                    local_up(st, &LS, bsf, (int)diag, st->lb_vector, &cnt_of_purn, alpha, diag_bias);
                }
                pruning_rate = cnt_of_purn / (double)(subcount - diag + 1);
                if (pruning_rate < 0.80) {
                    local_down_rounds++;
                    local_down(st, &LS, bsf, (int)diag, st->lb_vector, &cnt_of_purn, 0, diag_bias);
                }
                local_prune_cnt += (cnt_of_purn - cnt2);
            }
            for (int64_t i = 1; i <= subcount - diag + 1; i++) {
                all_cnt++;
                if (st->lb_vector[i - 1]) continue;
                kk_cnt++;
                double lbk1 = 0, lbk2 = 0;
                if (!lb_kk(SUB(st, i - 1), SUBU(st, i + diag - 2), SUBL(st, i + diag - 2), subseqLen, bsf * bsf, cb1,
                           st->special + 2 * (i - 1), &lbk1))
                    continue;
                if (!lb_kk(SUB(st, i + diag - 2), SUBU(st, i - 1), SUBL(st, i - 1), subseqLen, bsf * bsf, cb2,
                           st->special + 2 * (i - 1), &lbk2))
                    continue;
                /* :514-520 */
                pe_cnt++;
                if (!lb_petitjean(proj, proj_u, proj_l, SUB(st, i + diag - 2), SUBU(st, i + diag - 2),
                                  SUBL(st, i + diag - 2), SUB(st, i - 1), SUBU(st, i - 1), SUBL(st, i - 1), maxwarp,
                                  subseqLen, bsf * bsf, &lbk2, cb2))
                    continue;
                dtw_cnt++;
                int mm = subseqLen;
                if (lbk1 > lbk2) {
                    cb[mm - 1] = cb1[mm - 1];
                    for (int q = mm - 2; q >= 0; q--) cb[q] = cb[q + 1] + cb1[q];
                } else {
                    cb[mm - 1] = cb2[mm - 1];
                    for (int q = mm - 2; q >= 0; q--) cb[q] = cb[q + 1] + cb2[q];
                }
                double dtw_distance =
                    sqrt(orc_mon_dtw(SUB(st, i - 1), SUB(st, i + diag - 2), cb, subseqLen, maxwarp, bsf * bsf));
                if (dtw_distance <= bsf) {
                    first_min = (double)i;
                    sec_min = (double)(i + diag - 1);
                    bsf = dtw_distance;
                    found = 1;
                }
            }
        }
    }
    out->first_min = (int64_t)first_min;
    out->sec_min = (int64_t)sec_min;
    out->found_in_sweep = found;
    out->seed_i = seed_i < seed_j ? seed_i : seed_j;
    out->seed_j = seed_i < seed_j ? seed_j : seed_i;
    if (found) {
        out->i = out->first_min - 1;
        out->j = out->sec_min - 1;
    } else {
        out->i = out->seed_i;
        out->j = out->seed_j;
    }
    out->bsf = bsf;
    out->fast_cnt = fast_prune_cnt;
    out->global_cnt = global_prune_cnt;
    out->all_cnt = all_cnt;
    out->kk_cnt = kk_cnt;
    out->dtw_cnt = dtw_cnt;
    out->pe_cnt = pe_cnt;
    out->local_rounds = local_rounds;
    out->local_down_rounds = local_down_rounds;
    out->local_cnt = local_prune_cnt;
    out->bad_bin = LS.bad_bin;
    local_free(&LS);
    free(alpha_table);
    free(lb_vector_for_next);
    /* free everything */
    free(st->mu); free(st->sig); free(st->invsig); free(st->subs); free(st->subsU); free(st->subsL);
    free(min_index); free(cb_temp); free(st->UTS); free(st->LTS); free(st->dr_bwd); free(st->dc_bwd);
    free(st->dr_fwd); free(st->dc_fwd); free(sumU); free(sumL); free(sumU2); free(sumL2); free(tmpv);
    free(st->sumU_sumL); free(sumU2_sumL2); free(st->DUL2_fast); free(st->DUL_fast); free(nL); free(nU);
    free(st->normUL_trans); free(st->MASK); free(st->TS2); free(st->UTS_g); free(st->LTS_g);
    free(st->dr_bwd_g); free(st->dc_bwd_g); free(st->dr_fwd_g); free(st->dc_fwd_g); free(st->dr_bwdM);
    free(st->dr_fwdM); free(st->dc_bwdT2); free(st->dc_fwdT2);
    free(st->sumMASK_g); free(st->sumU_sumL_g); free(st->normUL_g); free(st->DUL2_g);
    free(st->DUL_g); free(st->special); free(st->lb_vector); free(st->lb_vector_new); free(cb); free(cb1);
    free(cb2); free(proj); free(proj_u); free(proj_l);
    return 0;
}

/* CPU_version1/DTWMotifDiscoveryGUI.cpp:12-27 */
int orc_motif_discovery(const double *a, int64_t n, int32_t m, int32_t w, orc_join_result *out) {
    int64_t subcount = n - m + 1;
    if (m < 8 || n < 2 * (int64_t)m + 1) return 1;
    double *mp = (double *)malloc(sizeof(double) * subcount);
    int32_t *mpi = (int32_t *)malloc(sizeof(int32_t) * subcount);
    int rc = orc_mpx(a, n, m, m, mp, mpi);
    if (rc == 0) rc = orc_join(a, n, m, w, mp, mpi, out);
    free(mp);
    free(mpi);
    return rc;
}

/* ---------------------------- independent oracle ------------------------------------ */
double orc_dtw_sq(const double *x, const double *y, int32_t m, int32_t w) {
    /* full (m+1)^2 table, cells outside |r-c| <= w stay infinite */
    size_t W = (size_t)m + 1;
    double *D = (double *)malloc(sizeof(double) * W * W);
    for (size_t k = 0; k < W * W; k++) D[k] = INFINITY;
    D[0] = 0.0;
    for (int r = 1; r <= m; r++) {
        int lo = r - w > 1 ? r - w : 1;
        int hi = r + w < m ? r + w : m;
        for (int c = lo; c <= hi; c++) {
            double d = x[r - 1] - y[c - 1];
            double best = D[(size_t)(r - 1) * W + (c - 1)];
            double up = D[(size_t)(r - 1) * W + c];
            double left = D[(size_t)r * W + (c - 1)];
            if (up < best) best = up;
            if (left < best) best = left;
            D[(size_t)r * W + c] = d * d + best;
        }
    }
    double res = D[(size_t)m * W + m];
    free(D);
    return res;
}

int orc_znorm(const double *a, int64_t n, int64_t i, int32_t m, double *z) {
    if (i < 0 || i + m > n) return 1;
    double s = 0.0;
    for (int k = 0; k < m; k++) s += a[i + k];
    double mu = s / m, v = 0.0;
    for (int k = 0; k < m; k++) v += (a[i + k] - mu) * (a[i + k] - mu);
    double sd = sqrt(v / m);
    if (!(sd > 0)) return 2;
    for (int k = 0; k < m; k++) z[k] = (a[i + k] - mu) / sd;
    return 0;
}

int orc_bruteforce_top1(const double *a, int64_t n, int32_t m, int32_t w, int64_t *bi, int64_t *bj, double *d) {
    int64_t subcount = n - m + 1;
    if (m < 2 || n < 2 * (int64_t)m) return 1;
    double *Z = (double *)malloc(sizeof(double) * (size_t)subcount * m);
    for (int64_t i = 0; i < subcount; i++)
        if (orc_znorm(a, n, i, m, Z + i * m)) {
            free(Z);
            return 2;
        }
    double best = INFINITY;
    int64_t best_i = -1, best_j = -1;
    for (int64_t i = 0; i < subcount; i++)
        for (int64_t j = i + m; j < subcount; j++) {
            double c = orc_dtw_sq(Z + i * m, Z + j * m, m, w);
            if (c < best) {
                best = c;
                best_i = i;
                best_j = j;
            }
        }
    free(Z);
    *bi = best_i;
    *bj = best_j;
    *d = sqrt(best);
    return 0;
}

/* LB_Keogh of window j against the radius-w envelope of window i over all m positions (no
 * abandoning), plus the LB_Petitjean increments of lb_petitjean above (q = window j, x = window
 * i: the orientation of the join's second LB_KK direction, CPU_version1/new_dtw_motifGUI.cpp:
 * 514-520), both windows z-normalised by orc_znorm, per-window envelopes by orc_lemire. */
double orc_lb_petitjean_pair(const double *a, int64_t n, int32_t m, int32_t w, int64_t i, int64_t j) {
    double *buf = (double *)xcalloc((size_t)m * 9, sizeof(double));
    double *zi = buf, *zj = buf + m, *ui = buf + 2 * m, *li = buf + 3 * m, *uj = buf + 4 * m, *lj = buf + 5 * m;
    double *p = buf + 6 * m, *up = buf + 7 * m, *lp = buf + 8 * m;
    double lbk = -1.0, cb_unused = 0.0;
    (void)cb_unused;
    if (orc_znorm(a, n, i, m, zi) == 0 && orc_znorm(a, n, j, m, zj) == 0) {
        orc_lemire(zi, m, w, li, ui);
        orc_lemire(zj, m, w, lj, uj);
        lbk = orc_lb_keogh_sq(zi, zj, m, w);
        double *cb = (double *)xcalloc((size_t)m, sizeof(double));
        lb_petitjean(p, up, lp, zj, uj, lj, zi, ui, li, w, m, ORC_INF, &lbk, cb);
        free(cb);
    }
    free(buf);
    return lbk;
}

double orc_lb_keogh_sq(const double *zi, const double *zj, int32_t m, int32_t w) {
    double s = 0.0;
    for (int k = 0; k < m; k++) {
        int lo = k - w > 0 ? k - w : 0, hi = k + w < m - 1 ? k + w : m - 1;
        double U = zi[lo], L = zi[lo];
        for (int t = lo; t <= hi; t++) {
            if (zi[t] > U) U = zi[t];
            if (zi[t] < L) L = zi[t];
        }
        if (zj[k] > U) s += (zj[k] - U) * (zj[k] - U);
        else if (zj[k] < L) s += (zj[k] - L) * (zj[k] - L);
    }
    return s;
}
