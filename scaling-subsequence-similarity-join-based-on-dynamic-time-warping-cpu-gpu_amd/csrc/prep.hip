// Per-series preprocessing on gfx950, all O(n) and HBM-bound:
//   * centring constant c0 = (min a + max a) / 2 and a non-finite check,
//   * radius-w lower/upper envelope U, L (reference: lower_upper_lemire,
//     CPU_version1/allmatrixop/lower_upper_lemire.cpp:7-43, clipped at the series ends),
//   * compensated FP64 (double-double) prefix sums of a-c0, (a-c0)^2, v-c0, (v-c0)^2, h^2 with
//     v = (U+L)/2, h = (U-L)/2, giving window means / population standard deviations
//     (reference: mvmean, CPU_version1/allmatrixop/movmean.cpp:9-33, which uses plain running
//     sums),
//   * the per-window table the diagonal sweep reads (WinRow), FP32 centred copies of the
//     series and envelope, and per-256-window maxima for the sweep's rounding-error margins.
#include "internal.hpp"

#include <cmath>

namespace dtwj {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 8;
constexpr int kScanBlock = kScanThreads * kScanPerThread;  // 2048 elements per scan block
constexpr int kRedBlocks = 1024;

__device__ inline unsigned int fbits(float x) { return __float_as_uint(x); }

__global__ void k_init(int32_t *flags, float *gstat) {
    int t = threadIdx.x;
    if (t < 4) flags[t] = 0;
    if (t < 8) gstat[t] = (t == 4) ? INFINITY : 0.0f;
}

__global__ __launch_bounds__(256) void k_minmax_partial(const double *__restrict__ a, int64_t n, double *red,
                                                        int32_t *flags) {
    __shared__ double smin[256], smax[256];
    double lo = INFINITY, hi = -INFINITY;
    int bad = 0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        double x = a[t];
        if (!isfinite(x)) bad = 1;
        lo = fmin(lo, x);
        hi = fmax(hi, x);
    }
    if (bad) atomicOr(&flags[0], 1);
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        red[2 * blockIdx.x] = smin[0];
        red[2 * blockIdx.x + 1] = smax[0];
    }
}

__global__ __launch_bounds__(256) void k_minmax_final(const double *red, int nparts, double *c0) {
    __shared__ double smin[256], smax[256];
    double lo = INFINITY, hi = -INFINITY;
    for (int p = threadIdx.x; p < nparts; p += 256) {
        lo = fmin(lo, red[2 * p]);
        hi = fmax(hi, red[2 * p + 1]);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *c0 = 0.5 * smin[0] + 0.5 * smax[0];
}

// Envelope: U[t] = max a[max(0,t-w) .. min(n-1,t+w)], L likewise; v = (U+L)/2.  The values are
// exact copies of series values (selection only).
//
// Sliding max / min by doubling in LDS, O(log w) per output with every lane busy: a tile of
// `tile` outputs stages its tile + 2w inputs once, with coalesced reads (positions outside
// the series read as -inf for the max and +inf for the min, which clips the windows at the series
// ends).  Pass k turns the max / min over runs of s = 2^k values into runs of 2s, in place
// (mx[i] <- max(mx[i], mx[i + s]); a chunk of positions reads before any of it is written and reads
// nothing a previous chunk wrote).  With s the largest power of two <= B = 2w + 1, the window of B
// values at i is the union of the runs at i and at i + B - s.
constexpr int kEnvThreads = 256;
// 256 outputs per tile while the halo 2w is small against it (more workgroups for short bands),
// 1024 beyond, so the halo stays under a fifth of the staged inputs up to w = 128.
static int envelope_tile(int w) { return w <= 16 ? 256 : 1024; }

__global__ __launch_bounds__(kEnvThreads) void k_envelope(const double *__restrict__ a, int64_t n, int w, int tile,
                                                          double *U, double *L, double *v) {
    extern __shared__ double env_lds[];
    const int B = 2 * w + 1;
    const int len = tile + 2 * w;
    double *mx = env_lds, *mn = env_lds + len;
    const int64_t t0 = (int64_t)blockIdx.x * tile;  // first output of the tile
    const int64_t base = t0 - w;                    // series position of buffer index 0
    for (int i = threadIdx.x; i < len; i += kEnvThreads) {
        const int64_t p = base + i;
        const bool in = p >= 0 && p < n;
        const double x = in ? a[p] : 0.0;
        mx[i] = in ? x : -INFINITY;
        mn[i] = in ? x : INFINITY;
    }
    __syncthreads();
    int s = 1;  // mx[i] = max of buffer [i, i + s) wherever i + s <= len
    for (; 2 * s <= B; s *= 2) {
        const int cnt = len - 2 * s + 1;  // positions whose run of 2s stays inside the buffer
        for (int c = 0; c < cnt; c += kEnvThreads) {
            const int i = c + threadIdx.x;
            double hi = 0.0, lo = 0.0;
            if (i < cnt) {
                hi = fmax(mx[i], mx[i + s]);
                lo = fmin(mn[i], mn[i + s]);
            }
            __syncthreads();
            if (i < cnt) {
                mx[i] = hi;
                mn[i] = lo;
            }
            __syncthreads();
        }
    }
    const int off = B - s;  // 0 <= off < s, and tile - 1 + off + s <= len
    for (int i = threadIdx.x; i < tile; i += kEnvThreads) {
        const int64_t t = t0 + i;
        if (t >= n) break;
        const double hi = fmax(mx[i], mx[i + off]);
        const double lo = fmin(mn[i], mn[i + off]);
        U[t] = hi;
        L[t] = lo;
        v[t] = 0.5 * hi + 0.5 * lo;
    }
}

// Channel values of the prefix-sum scan, squares as exact double-double products (TwoProd).
__device__ inline dd_t exact_sq(double x) {
    const double p = x * x;
    return {p, fma(x, x, -p)};
}
__device__ inline void scan_values(const double *__restrict__ a, const double *__restrict__ U,
                                   const double *__restrict__ L, double c0, int64_t t, dd_t x[kPrepScanChannels]) {
    double ad = a[t] - c0;
    double u = U[t], l = L[t];
    double vd = (0.5 * u + 0.5 * l) - c0;
    double h = 0.5 * u - 0.5 * l;
    x[0] = {ad, 0.0};
    x[1] = exact_sq(ad);
    x[2] = {vd, 0.0};
    x[3] = exact_sq(vd);
    x[4] = exact_sq(h);
    x[5] = {h, 0.0};
}

// Inclusive block scan of one dd value per thread (Hillis-Steele in LDS).
__device__ inline dd_t block_scan_dd(dd_t x, dd_t *buf) {
    int t = threadIdx.x;
    buf[t] = x;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        dd_t y = buf[t];
        if (t >= off) y = dd_add(buf[t - off], y);
        __syncthreads();
        buf[t] = y;
        __syncthreads();
    }
    dd_t r = buf[t];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_partial(const double *__restrict__ a,
                                                               const double *__restrict__ U,
                                                               const double *__restrict__ L, const double *c0p,
                                                               int64_t n, int64_t nb, dd_t *bsum) {
    __shared__ dd_t buf[kScanThreads];
    const double c0 = *c0p;
    dd_t acc[kPrepScanChannels];
    for (int c = 0; c < kPrepScanChannels; ++c) acc[c] = {0.0, 0.0};
    int64_t base = (int64_t)blockIdx.x * kScanBlock + (int64_t)threadIdx.x * kScanPerThread;
    for (int k = 0; k < kScanPerThread; ++k) {
        int64_t t = base + k;
        if (t < n) {
            dd_t x[kPrepScanChannels];
            scan_values(a, U, L, c0, t, x);
            for (int c = 0; c < kPrepScanChannels; ++c) acc[c] = dd_add(acc[c], x[c]);
        }
    }
    for (int c = 0; c < kPrepScanChannels; ++c) {
        dd_t s = block_scan_dd(acc[c], buf);
        if (threadIdx.x == kScanThreads - 1) bsum[(int64_t)c * nb + blockIdx.x] = s;
    }
}

// Exclusive scan of the block sums, one block per channel.
__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(dd_t *bsum, int64_t nb) {
    __shared__ dd_t buf[kScanThreads];
    dd_t *b = bsum + (int64_t)blockIdx.x * nb;
    int64_t per = (nb + kScanThreads - 1) / kScanThreads;
    int64_t beg = (int64_t)threadIdx.x * per;
    int64_t end = beg + per < nb ? beg + per : nb;
    dd_t s = {0.0, 0.0};
    for (int64_t k = beg; k < end; ++k) s = dd_add(s, b[k]);
    dd_t inc = block_scan_dd(s, buf);
    dd_t run = dd_sub(inc, s);  // exclusive prefix of this thread's chunk
    for (int64_t k = beg; k < end; ++k) {
        dd_t x = b[k];
        b[k] = run;
        run = dd_add(run, x);
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_final(const double *__restrict__ a, const double *__restrict__ U,
                                                             const double *__restrict__ L, const double *c0p,
                                                             int64_t n, int64_t nb, const dd_t *bsum, dd_t *P) {
    __shared__ dd_t buf[kScanThreads];
    const double c0 = *c0p;
    int64_t base = (int64_t)blockIdx.x * kScanBlock + (int64_t)threadIdx.x * kScanPerThread;
    dd_t xs[kScanPerThread][kPrepScanChannels];
    dd_t tot[kPrepScanChannels];
    for (int c = 0; c < kPrepScanChannels; ++c) tot[c] = {0.0, 0.0};
    for (int k = 0; k < kScanPerThread; ++k) {
        int64_t t = base + k;
        if (t < n) {
            scan_values(a, U, L, c0, t, xs[k]);
        } else {
            for (int c = 0; c < kPrepScanChannels; ++c) xs[k][c] = {0.0, 0.0};
        }
        for (int c = 0; c < kPrepScanChannels; ++c) tot[c] = dd_add(tot[c], xs[k][c]);
    }
    for (int c = 0; c < kPrepScanChannels; ++c) {
        dd_t inc = block_scan_dd(tot[c], buf);
        dd_t run = dd_add(bsum[(int64_t)c * nb + blockIdx.x], dd_sub(inc, tot[c]));
        dd_t *Pc = P + (int64_t)c * (n + 1);
        if (blockIdx.x == 0 && threadIdx.x == 0) Pc[0] = {0.0, 0.0};
        for (int k = 0; k < kScanPerThread; ++k) {
            int64_t t = base + k;
            run = dd_add(run, xs[k][c]);
            if (t < n) Pc[t + 1] = run;
        }
    }
}

struct WinStats {
    double mu, sig, muv, sv, N, R;
    bool degenerate;
};

__device__ inline WinStats window_stats(const dd_t *P, int64_t n, int m, double c0, int64_t i) {
    dd_t S[5];
    for (int c = 0; c < 5; ++c) S[c] = dd_sub(P[(int64_t)c * (n + 1) + i + m], P[(int64_t)c * (n + 1) + i]);
    WinStats r;
    dd_t mean_d = dd_div_d(S[0], (double)m);
    r.mu = c0 + (mean_d.hi + mean_d.lo);
    // m * var = S1 - S0^2 / m, in double-double
    dd_t mv = dd_sub(S[1], dd_div_d(dd_mul(S[0], S[0]), (double)m));
    double var = (mv.hi + mv.lo) / m;
    double mabs = fabs(mean_d.hi) + 1.0;
    r.degenerate = !(var > 0.0) || sqrt(var) <= 0x1p-36 * mabs;
    r.sig = sqrt(var > 0.0 ? var : 0.0);
    dd_t meanv_d = dd_div_d(S[2], (double)m);
    r.muv = c0 + (meanv_d.hi + meanv_d.lo);
    dd_t vv = dd_sub(S[3], dd_div_d(dd_mul(S[2], S[2]), (double)m));  // sum (v - muv)^2
    double svv = vv.hi + vv.lo;
    if (svv < 0.0) svv = 0.0;
    r.sv = sqrt(svv / m);
    double dmu = (meanv_d.hi + meanv_d.lo) - (mean_d.hi + mean_d.lo);  // muv - mu, from the centred means
    double var_safe = r.sig > 0.0 ? r.sig * r.sig : 1.0;
    r.N = (svv + (double)m * dmu * dmu) / var_safe;
    double h2 = S[4].hi + S[4].lo;
    r.R = sqrt(h2 > 0.0 ? h2 : 0.0) / (r.sig > 0.0 ? r.sig : 1.0);
    return r;
}

// Bits of the smallest FP16 value >= v (v >= 0; +inf beyond the FP16 range).
__device__ inline uint32_t f16_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, f) + 1u);
    const _Float16 h = (_Float16)f;
    uint32_t b = (uint32_t)__builtin_bit_cast(uint16_t, h);
    if ((float)h < f) ++b;  // h is finite here (an infinite h is above f)
    return b;
}

__device__ inline void atomic_max_pos(float *p, float x) { atomicMax((unsigned int *)p, fbits(x)); }
__device__ inline void atomic_min_pos(float *p, float x) { atomicMin((unsigned int *)p, fbits(x)); }

// One thread per window; each block of 256 windows also writes its maxima (sigma, sv, 1/sigma).
__global__ __launch_bounds__(256) void k_window_stats(const double *__restrict__ a, const double *__restrict__ v,
                                                      const dd_t *__restrict__ P, const double *c0p, int64_t n,
                                                      int m, double *mu, double *sig, double *muv, float *mu_c,
                                                      float *isg, WinRow *win, float *blkmax, int nblk, float *gstat,
                                                      int32_t *flags) {
    __shared__ float r0[256], r1[256], r2[256], r3[256], r4[256];
    const double c0 = *c0p;
    const int64_t nw = n - m + 1;
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float fs = 0.f, fsv = 0.f, fis = 0.f, fmc = 0.f, fsmin = __builtin_huge_valf();
    if (i < nw) {
        WinStats s = window_stats(P, n, m, c0, i);
        if (s.degenerate) atomicOr(&flags[1], 1);
        mu[i] = s.mu;
        sig[i] = s.sig;
        muv[i] = s.muv;
        WinRow row;
        if (i + 1 < nw) {
            WinStats s1 = window_stats(P, n, m, c0, i + 1);
            double am = a[i + m], a0 = a[i], vm = v[i + m], v0 = v[i];
            row.dfa = (float)(0.5 * (am - a0));
            row.dga = (float)((am - s1.mu) + (a0 - s.mu));
            row.dfv = (float)(0.5 * (vm - v0));
            row.dgv = (float)((vm - s1.muv) + (v0 - s.muv));
        } else {
            row.dfa = row.dga = row.dfv = row.dgv = 0.f;
        }
        double is = s.sig > 0.0 ? 1.0 / s.sig : 0.0;
        row.isig = (float)is;
        isg[i] = row.isig;
        row.mN = (float)((double)m + s.N);
        row.R = (float)s.R;
        row.sig = (float)s.sig;
        win[i] = row;
        float mc = (float)(s.mu - c0);
        mu_c[i] = mc;
        fs = (float)s.sig;
        fsv = (float)s.sv;
        fis = (float)is;
        fmc = fabsf(mc);
        fsmin = fs;
    }
    // block maxima (and the minimum of sigma), then one set of atomics per block for the
    // series-wide statistics
    r0[threadIdx.x] = fs;
    r1[threadIdx.x] = fsv;
    r2[threadIdx.x] = fis;
    r3[threadIdx.x] = fmc;
    r4[threadIdx.x] = fsmin;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            r0[threadIdx.x] = fmaxf(r0[threadIdx.x], r0[threadIdx.x + s]);
            r1[threadIdx.x] = fmaxf(r1[threadIdx.x], r1[threadIdx.x + s]);
            r2[threadIdx.x] = fmaxf(r2[threadIdx.x], r2[threadIdx.x + s]);
            r3[threadIdx.x] = fmaxf(r3[threadIdx.x], r3[threadIdx.x + s]);
            r4[threadIdx.x] = fminf(r4[threadIdx.x], r4[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if ((int)blockIdx.x < nblk) {
            // round maxima up by one part in 2^20 so the FP32 conversion never underestimates
            blkmax[0 * nblk + blockIdx.x] = r0[0] * (1.0f + 0x1p-20f);
            blkmax[1 * nblk + blockIdx.x] = r1[0] * (1.0f + 0x1p-20f);
            blkmax[2 * nblk + blockIdx.x] = r2[0] * (1.0f + 0x1p-20f);
        }
        atomic_max_pos(&gstat[1], r3[0]);
        atomic_max_pos(&gstat[2], r2[0]);
        atomic_max_pos(&gstat[3], r0[0]);
        atomic_min_pos(&gstat[4], r4[0]);  // every block holds at least one window
    }
}

__global__ __launch_bounds__(256) void k_positions(const double *__restrict__ a, const double *__restrict__ U,
                                                   const double *__restrict__ L, const double *c0p, int64_t n,
                                                   float *ac, float *uc, float *lc, uint32_t *wc, float *gstat) {
    __shared__ float red[256];
    const double c0 = *c0p;
    int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float mx = 0.f;
    if (t < n) {
        float x = (float)(a[t] - c0), u = (float)(U[t] - c0), l = (float)(L[t] - c0);
        ac[t] = x;
        uc[t] = u;
        lc[t] = l;
        // widths of the FP32 envelope around the FP32 sample, exact in FP64, as the smallest
        // FP16 values not below them: x + wU >= u and x - wL <= l hold after any FP32 rounding
        wc[t] = f16_up((double)u - (double)x) | (f16_up((double)x - (double)l) << 16);
        mx = fmaxf(fabsf(x), fmaxf(fabsf(u), fabsf(l)));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomic_max_pos(&gstat[0], red[0]);
}

}  // namespace

int64_t prep_scan_blocks(int64_t n) { return (n + kScanBlock - 1) / kScanBlock; }

// Segment-mean table (layout at paa_segments): one thread per (window, segment), segments
// fastest so that the stores of a block are contiguous.  The segment means are differences of
// the compensated prefix sums (channels a - c0, v - c0 and h), taken against the window mean in
// double-double before the division by sigma, so a large offset of the series costs no accuracy.
// |Z| <= sqrt(m) (a mean of z-values), so Z rounded to FP32 is within 2^-24 sqrt(m) of its FP64
// value; U and L are widened by twice that (which also covers the FP64 rounding here) and then
// rounded outwards, so that the FP32 table bounds every window pair's exact segment terms.
__global__ __launch_bounds__(256) void k_paa_tables(const dd_t *__restrict__ P, int64_t n, int m, int nseg,
                                                    const double *__restrict__ sig, float *paa, int64_t nw) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nw * nseg) return;
    const int64_t p = t / nseg;
    const int sg = (int)(t - p * nseg);
    const int len = m / nseg;
    const dd_t *P0 = P, *P2 = P + 2 * (n + 1), *P5 = P + 5 * (n + 1);
    const dd_t mean = dd_div_d(dd_sub(P0[p + m], P0[p]), (double)m);
    const int64_t b = p + (int64_t)sg * len;
    const dd_t za = dd_sub(dd_div_d(dd_sub(P0[b + len], P0[b]), (double)len), mean);
    const dd_t zv = dd_sub(dd_div_d(dd_sub(P2[b + len], P2[b]), (double)len), mean);
    const dd_t hh = dd_div_d(dd_sub(P5[b + len], P5[b]), (double)len);
    const double is = 1.0 / sig[p];
    const double z = (za.hi + za.lo) * is;
    const double mv = (zv.hi + zv.lo) * is;
    const double h = (hh.hi + hh.lo) * is;
    const double e = 0x1p-23 * sqrt((double)m);
    float *row = paa + p * 3 * nseg;
    row[sg] = __double2float_ru(mv + h + e);
    row[nseg + sg] = __double2float_rd(mv - h - e);
    row[2 * nseg + sg] = (float)z;
}

int prep_series(Prep &P, hipStream_t s) {
    const int64_t n = P.n;
    const int m = P.m;
    const int64_t nw = n - m + 1;
    const int64_t nb = prep_scan_blocks(n);
    if (P.w < 0 || P.w > kMaxEnvelopeRadius) return hip_fail(hipErrorInvalidValue, "envelope radius", __FILE__, __LINE__);
    hipLaunchKernelGGL(k_init, dim3(1), dim3(64), 0, s, P.flags, P.gstat);
    int parts = (int)((n + 255) / 256 < kRedBlocks ? (n + 255) / 256 : kRedBlocks);
    hipLaunchKernelGGL(k_minmax_partial, dim3(parts), dim3(256), 0, s, P.a, n, P.red, P.flags);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, s, P.red, parts, P.c0);
    const int etile = envelope_tile(P.w);
    hipLaunchKernelGGL(k_envelope, dim3((unsigned)((n + etile - 1) / etile)), dim3(kEnvThreads),
                       sizeof(double) * 2 * (size_t)(etile + 2 * P.w), s, P.a, n, P.w, etile, P.U, P.L, P.v);
    hipLaunchKernelGGL(k_scan_partial, dim3((unsigned)nb), dim3(kScanThreads), 0, s, P.a, P.U, P.L, P.c0, n, nb,
                       P.bsum);
    hipLaunchKernelGGL(k_scan_blocks, dim3(kPrepScanChannels), dim3(kScanThreads), 0, s, P.bsum, nb);
    hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nb), dim3(kScanThreads), 0, s, P.a, P.U, P.L, P.c0, n, nb, P.bsum,
                       P.P);
    hipLaunchKernelGGL(k_window_stats, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, P.a, P.v, P.P, P.c0, n, m,
                       P.mu, P.sig, P.muv, P.mu_c, P.isg, P.win, P.blkmax, P.nblk, P.gstat, P.flags);
    hipLaunchKernelGGL(k_positions, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P.a, P.U, P.L, P.c0, n, P.ac,
                       P.uc, P.lc, P.wc, P.gstat);
    if (P.paa) {
        if ((P.paa_seg != 16 && P.paa_seg != 32) || m < P.paa_seg)
            return hip_fail(hipErrorInvalidValue, "segment-mean table", __FILE__, __LINE__);
        const int64_t cells = nw * P.paa_seg;
        hipLaunchKernelGGL(k_paa_tables, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, P.P, n, m, P.paa_seg,
                           P.sig, P.paa, nw);
    }
    DTWJ_HIP(hipGetLastError());
    return 0;
}

}  // namespace dtwj
