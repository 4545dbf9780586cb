// C ABI of the gfx950 top-1 banded-DTW join (include/dtwj.h): argument validation, device
// workspace, the launch sequence prep -> seed -> fused join -> FP64 re-verification, and the
// reference-style reporting.  Host code only; kernels live in prep.hip / join.hip / dtw.hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "../../include/dtwj.h"
#include "internal.hpp"

namespace dtwj {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file,
                  line);
    set_error(buf);
    return e == hipErrorOutOfMemory ? DTWJ_ENOMEM : DTWJ_EDEVICE;
}

namespace {

using clk = std::chrono::steady_clock;
inline double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

constexpr int64_t kCandCap = 1 << 20;       // candidate slots of a single-series join
constexpr int64_t kBatchCandCap = 1 << 14;  // ... per series of a batch
constexpr int64_t kMaxCandCap = 0x7fffffffLL;  // largest regrown buffer (re-scoring grid limit)
constexpr int kRegrow = -2;  // finish_job: more candidates were pushed than the buffer holds
constexpr int kDiagGroup = 256;  // diagonals per work item (join.hip DG)
constexpr int kNCounters = 9;

int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}

// A bump allocator over one device allocation, 256-byte aligned slices.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, off = 0;
    template <typename T>
    T *take(size_t count) {
        size_t bytes = (count * sizeof(T) + 255) & ~(size_t)255;
        T *p = reinterpret_cast<T *>(base + off);
        off += bytes;
        return p;
    }
};

}  // namespace

}  // namespace dtwj

using namespace dtwj;

struct dtwj_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void *mem = nullptr;
    size_t mem_cap = 0;
    double *upload = nullptr;
    int64_t upload_cap = 0;
    int num_cus = 0;
    // workspace of fixed size that must not be reallocated (the no-allocation entry point's
    // per-device array; no events either: the join is timed on the host clock)
    bool fixed = false;
    // batches: tile choices of the autotuner, keyed by the batch's shape
    std::map<std::string, std::pair<int32_t, int32_t>> tuned;
};

extern "C" {

void dtwj_default_options(dtwj_options *opt) {
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->device = 0;
}

int32_t dtwj_band_from_ratio(int32_t m, double r) {
    // GPU_version1/src/AAAA.cu:22-34 parses r as float and computes MAX(round(m*w), 1)
    float prod = (float)m * (float)r;
    int32_t w = (int32_t)std::lround((double)prod);
    return w > 1 ? w : 1;
}

const char *dtwj_last_error(void) { return g_last_error.c_str(); }

int32_t dtwj_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return c;
}

void dtwj_free(void *p) { std::free(p); }

int dtwj_context_create(int32_t device, dtwj_context **out) {
    if (!out) return fail(DTWJ_EINVAL, "null context pointer");
    *out = nullptr;
    int count = dtwj_device_count();
    if (count <= 0) return fail(DTWJ_EDEVICE, "no HIP device available");
    if (device < 0 || device >= count) return fail(DTWJ_EINVAL, "device ordinal out of range");
    dtwj_context *c = new dtwj_context();
    c->device = device;
    hipDeviceProp_t prop;
    hipError_t e = hipSetDevice(device);
    const char *what = "hipSetDevice";
    if (e == hipSuccess) {
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        what = "hipStreamCreateWithFlags";
    }
    if (e == hipSuccess) {
        e = hipEventCreate(&c->ev0);
        what = "hipEventCreate";
    }
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) {
        e = hipGetDeviceProperties(&prop, device);
        what = "hipGetDeviceProperties";
    }
    if (e != hipSuccess) {
        const int rc = hip_fail(e, what, __FILE__, __LINE__);
        dtwj_context_destroy(c);  // releases whatever was created
        return rc;
    }
    c->num_cus = prop.multiProcessorCount;
    *out = c;
    return DTWJ_OK;
}

void dtwj_context_destroy(dtwj_context *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->mem) (void)hipFree(c->mem);
    if (c->upload) (void)hipFree(c->upload);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int dtwj_context_upload(dtwj_context *c, const double *series, int64_t n, const double **d_series) {
    if (!c || !series || n <= 0 || !d_series) return fail(DTWJ_EINVAL, "bad upload arguments");
    DTWJ_HIP(hipSetDevice(c->device));
    if (n > c->upload_cap) {
        if (c->upload) DTWJ_HIP(hipFree(c->upload));
        c->upload = nullptr;
        DTWJ_HIP(hipMalloc(&c->upload, sizeof(double) * n));
        c->upload_cap = n;
    }
    DTWJ_HIP(hipMemcpyAsync(c->upload, series, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    DTWJ_HIP(hipStreamSynchronize(c->stream));
    *d_series = c->upload;
    return DTWJ_OK;
}

}  // extern "C"

namespace {

int validate(const double *series, int64_t n, int32_t m, int32_t w) {
    if (!series) return fail(DTWJ_EINVAL, "series is NULL");
    if (m < 8) return fail(DTWJ_EINVAL, "subsequence length m must be >= 8");
    if (w < 0) return fail(DTWJ_EINVAL, "band radius w must be >= 0");
    if (n < 2 * (int64_t)m) return fail(DTWJ_EINVAL, "series too short: need n >= 2m");
    if (n > ((int64_t)1 << 31) - 1) return fail(DTWJ_EINVAL, "series too long: n must be < 2^31");
    const int32_t we = std::min(w, m - 1);
    if (2 * we + 1 > 1024) return fail(DTWJ_EINVAL, "band too wide: 2*min(w, m-1)+1 must be <= 1024");
    return DTWJ_OK;
}

// ---------------------------------------------------------------- one series of a join
// Sizes, device slices and the host copy of the join descriptor of one series.
struct Job {
    const double *d_a = nullptr;
    int64_t n = 0, nw = 0, K = 0;
    int32_t m = 0, w = 0;  // w already clamped to m - 1
    int32_t groups = 0, rows_per_item = 0, n_seed = 0;
    int64_t cand_cap = 0;
    int64_t cand_need = 0;   // pushes counted by the last run when they exceeded cand_cap
    unsigned long long init_key = 0;  // best-so-far the join starts from
    int32_t sweep_mode = 0;  // dtwj_options::sweep
    int32_t paa_seg = 0;     // segments per window of the segment-mean stage (0: stage off)
    int64_t n_items = 0;
    std::vector<int64_t> item_start;
    std::vector<int2> extra;  // explicit seed pairs (besides the scan's and (0, m))
    Prep P;
    JoinArgs A;               // host copy: pointers and sizes (error model filled on the device)
    unsigned long long *d_bsf = nullptr, *d_counters = nullptr;
    int4 *d_cands = nullptr;
    unsigned long long *d_ncands = nullptr;
    int64_t *d_item_start = nullptr;
    int2 *d_seed_pairs = nullptr;
    int32_t n_seed_pairs = 0;
};

// Sweep tile height: a multiple of 128 rows, at most 2^20 (queue entries carry the row
// offset inside the item in 22 bits).  Automatic: 3072 rows for bands up to 3 (survivors are
// rare there, so a taller item amortises its FP64 anchor at little cost in the FP32 drift
// margin), 2048 otherwise.
int32_t normalise_rows(int32_t r, int32_t w) {
    r = r > 0 ? std::min(r, 1 << 20) : (w <= 3 ? 3072 : 2048);
    return std::max(128, (r + 127) / 128 * 128);
}

void plan_job(Job &J, const double *d_a, int64_t n, int32_t m, int32_t w_in, int32_t rows_per_item, int32_t seed_req,
              bool gpu_scan, const std::vector<int2> &extra, int64_t cand_cap) {
    J.d_a = d_a;
    J.n = n;
    J.m = m;
    J.w = std::min(w_in, m - 1);
    J.nw = n - m + 1;
    J.K = n - 2 * (int64_t)m + 1;  // diagonal offsets m .. n-m
    J.groups = (int32_t)((J.K + kDiagGroup - 1) / kDiagGroup);
    J.rows_per_item = normalise_rows(rows_per_item, J.w);
    J.item_start.assign(J.groups + 1, 0);
    for (int32_t g = 0; g < J.groups; ++g) {
        const int64_t rows_g = J.nw - ((int64_t)m + (int64_t)kDiagGroup * g);
        J.item_start[g + 1] = J.item_start[g] + (rows_g + J.rows_per_item - 1) / J.rows_per_item;
    }
    J.n_items = J.item_start[J.groups];
    // 1024 sampled items: the join's time barely depends on the seed (its own best-so-far falls
    // within the first chunks), so a larger scan costs more than it saves
    int32_t ns = seed_req > 0 ? seed_req : 1024;
    if (!gpu_scan) ns = 0;
    if (ns > J.n_items) ns = (int32_t)J.n_items;
    J.n_seed = ns;
    J.extra = extra;
    J.n_seed_pairs = ns + 1 + (int32_t)extra.size();
    J.cand_cap = cand_cap;
    J.init_key = pack_bsf(3.0e38f, 0xffffffffu);
}

// Segments per window of the segment-mean stage for an effective band w (0: the stage is off).
// Automatic (mode 0): bands of 8 and more with windows of 128 and more, where LB_Keogh takes a
// visible share of the join; mode 1 forces it on (windows of 16 and more), mode 2 off.  The
// stage only prunes, so the answer is the same either way.
int32_t paa_seg_for(int32_t m, int32_t w, int32_t mode) {
    if (mode == 2 || m < 16) return 0;
    if (mode == 0 && (w < 8 || m < 128)) return 0;
    return paa_segments(m);
}

// Device bytes of one series (statistics, envelopes, tables, scans, join state).
size_t job_bytes(const Job &J) {
    const int64_t n = J.n, nw = J.nw;
    const int64_t nb = prep_scan_blocks(n);
    const int64_t nblk = (nw + 255) / 256;
    size_t b = 0;
    auto add = [&](size_t bytes) { b += (bytes + 255) & ~(size_t)255; };
    add(8 * n * 3);       // U L v
    add(8 * nw * 3);      // mu sig muv
    add(8);               // c0
    add(4 * n * 3);       // ac uc lc
    add(4 * n);           // wc
    add(4 * nw);          // mu_c
    add(4 * nw);          // isg
    add(sizeof(WinRow) * nw);
    add(4 * 3 * nblk);
    add(4 * 8);           // gstat
    add(sizeof(dd_t) * kPrepScanChannels * (n + 1));
    add(sizeof(dd_t) * kPrepScanChannels * nb);
    if (J.paa_seg > 0) add(sizeof(float) * 3 * (size_t)J.paa_seg * (size_t)nw);
    add(8 * 2 * 1024);
    add(4 * 4);           // flags
    add(8);               // bsf
    add(8 * (J.groups + 1));
    add(16 * (size_t)J.cand_cap);
    add(8);               // n_cands
    add(8 * kNCounters);
    add(8 * ((size_t)J.n_seed_pairs + 1));
    return b;
}

void take_job(Arena &ar, Job &J) {
    Prep &P = J.P;
    const int64_t n = J.n, nw = J.nw;
    P.a = J.d_a;
    P.n = n;
    P.m = J.m;
    P.w = J.w;
    P.U = ar.take<double>(n);
    P.L = ar.take<double>(n);
    P.v = ar.take<double>(n);
    P.mu = ar.take<double>(nw);
    P.sig = ar.take<double>(nw);
    P.muv = ar.take<double>(nw);
    P.c0 = ar.take<double>(1);
    P.ac = ar.take<float>(n);
    P.uc = ar.take<float>(n);
    P.lc = ar.take<float>(n);
    P.wc = ar.take<uint32_t>(n);
    P.mu_c = ar.take<float>(nw);
    P.isg = ar.take<float>(nw);
    P.win = ar.take<WinRow>(nw);
    P.nblk = (int32_t)((nw + 255) / 256);
    P.blkmax = ar.take<float>(3 * (size_t)P.nblk);
    P.gstat = ar.take<float>(8);
    P.P = ar.take<dd_t>(kPrepScanChannels * (size_t)(n + 1));
    P.bsum = ar.take<dd_t>(kPrepScanChannels * (size_t)prep_scan_blocks(n));
    P.paa = J.paa_seg > 0 ? ar.take<float>(3 * (size_t)J.paa_seg * (size_t)nw) : nullptr;
    P.paa_seg = J.paa_seg;
    P.red = ar.take<double>(2 * 1024);
    P.flags = ar.take<int32_t>(4);
    J.d_bsf = ar.take<unsigned long long>(1);
    J.d_item_start = ar.take<int64_t>(J.groups + 1);
    J.d_cands = ar.take<int4>(J.cand_cap);
    J.d_ncands = ar.take<unsigned long long>(1);
    J.d_counters = ar.take<unsigned long long>(kNCounters);
    J.d_seed_pairs = ar.take<int2>((size_t)J.n_seed_pairs + 1);

    JoinArgs &A = J.A;
    std::memset(&A, 0, sizeof(A));
    A.a = J.d_a;
    A.v = P.v;
    A.mu = P.mu;
    A.muv = P.muv;
    A.win = P.win;
    A.ac = P.ac;
    A.uc = P.uc;
    A.lc = P.lc;
    A.wc = P.wc;
    A.mu_c = P.mu_c;
    A.isg = P.isg;
    A.blkmax = P.blkmax;
    A.gstat = P.gstat;
    A.flags = P.flags;
    A.nblk = P.nblk;
    A.n = n;
    A.nw = nw;
    A.m = J.m;
    A.w = J.w;
    A.item_start = J.d_item_start;
    A.groups = J.groups;
    A.rows_per_item = J.rows_per_item;
    A.n_items = J.n_items;
    A.bsf = J.d_bsf;
    A.init_key = J.init_key;
    A.cands = J.d_cands;
    A.n_cands = J.d_ncands;
    A.cand_cap = J.cand_cap;
    A.counters = J.d_counters;
    A.sweep_mode = J.sweep_mode;
    A.paa = P.paa;
    A.paa_seg = P.paa_seg;
    A.paa_len = P.paa_seg > 0 ? J.m / P.paa_seg : 0;
}

// Host-to-device copies a job needs before its kernels run (item offsets, explicit seeds).
int upload_job(Job &J, hipStream_t s) {
    DTWJ_HIP(hipMemcpyAsync(J.d_item_start, J.item_start.data(), sizeof(int64_t) * (J.groups + 1),
                            hipMemcpyHostToDevice, s));
    std::vector<int2> fixed = J.extra;
    if (J.n_seed == 0) fixed.insert(fixed.begin(), make_int2(0, J.m));  // else written by the scan
    if (!fixed.empty()) {
        const int32_t off = J.n_seed == 0 ? 0 : J.n_seed + 1;
        DTWJ_HIP(hipMemcpyAsync(J.d_seed_pairs + off, fixed.data(), sizeof(int2) * fixed.size(),
                                hipMemcpyHostToDevice, s));
    }
    return 0;
}

// Seed scan (ED over sampled items) + DTW of every seed pair.  d_args: the job's descriptor.
// Seed-scan arguments of a job whose device buffers are placed.
SeedArgs seed_args(const Job &J) {
    SeedArgs S;
    S.a = J.d_a;
    S.mu = J.P.mu;
    S.win = J.P.win;
    S.n = J.n;
    S.nw = J.nw;
    S.m = J.m;
    S.item_start = J.d_item_start;
    S.groups = J.groups;
    S.n_items = J.n_items;
    S.rows_per_item = J.rows_per_item;
    S.rows = std::min(J.rows_per_item, 512);
    S.n_seed = J.n_seed;
    return S;
}

int enqueue_seeds(Job &J, const JoinArgs *d_args, hipStream_t s) {
    if (J.n_seed > 0)
        if (int rc = launch_seed(seed_args(J), J.d_seed_pairs, s)) return rc;
    return launch_seed_dtw(d_args, J.m, J.w, J.d_seed_pairs, J.n_seed_pairs, s);
}

// Reads the job's join state back, re-scores on the device every candidate within the FP32
// error of the best in FP64 and fills *out.  d_cost64: scratch for cand_cap doubles; d_part /
// d_win: rescore_partials() + 1 Winner slots.  Returns kRegrow (with J.cand_need set) when the
// join pushed more candidates than its buffer holds: the caller re-runs it with a larger buffer.
int finish_job(Job &J, double *d_cost64, Winner *d_part, Winner *d_win, hipStream_t s, dtwj_result *out) {
    unsigned long long ncands = 0, key = 0, counters[kNCounters];
    DTWJ_HIP(hipMemcpyAsync(&ncands, J.d_ncands, sizeof(ncands), hipMemcpyDeviceToHost, s));
    DTWJ_HIP(hipMemcpyAsync(&key, J.d_bsf, sizeof(key), hipMemcpyDeviceToHost, s));
    DTWJ_HIP(hipMemcpyAsync(counters, J.d_counters, sizeof(counters), hipMemcpyDeviceToHost, s));
    DTWJ_HIP(hipStreamSynchronize(s));
    if (ncands > (unsigned long long)J.cand_cap) {
        J.cand_need = (int64_t)std::min<unsigned long long>(ncands, (unsigned long long)INT64_MAX);
        J.init_key = key;
        return kRegrow;
    }
    if (ncands == 0) return fail(DTWJ_EDEVICE, "join produced no candidate");
    const float best32 = bsf_cost(key);
    const double eps = J.A.eps_z, rel = J.A.rel_dtw;
    const int m = J.m;
    auto err = [&](double D) { return rel * D + 4.0 * eps * std::sqrt(2.0 * m * D) + 8.0 * m * eps * eps; };
    const double T2 = (double)best32 + err(best32);
    const float ub = (float)(T2 + err(T2));  // rounded to nearest: the test below keeps a margin
    const float ub_up = std::nextafter(ub, std::numeric_limits<float>::infinity());
    if (int rc = launch_rescore(J.A, J.P.sig, J.d_cands, (int64_t)ncands, ub_up, d_cost64, d_part, d_win, s)) return rc;
    Winner win;
    DTWJ_HIP(hipMemcpyAsync(&win, d_win, sizeof(win), hipMemcpyDeviceToHost, s));
    DTWJ_HIP(hipStreamSynchronize(s));
    if (win.verified == 0 || !std::isfinite(win.cost)) return fail(DTWJ_EDEVICE, "no candidate re-scored in FP64");
    std::memset(out, 0, sizeof(*out));
    out->i = win.i;
    out->j = win.j;
    out->dist = std::sqrt(win.cost);
    out->dist_fp32 = std::sqrt((double)win.cost32);
    out->pairs_total = J.K * (J.K + 1) / 2;
    out->pairs_swept = (int64_t)counters[0];
    out->pass_ball = (int64_t)counters[1];
    out->pass_kim = (int64_t)counters[2];
    out->pass_keogh = (int64_t)counters[3];
    out->pass_paa = J.paa_seg > 0 ? (int64_t)counters[8] : (int64_t)counters[2];
    out->dtw_abandoned = (int64_t)counters[4];
    out->dtw_cells = (int64_t)counters[5];
    out->keogh_elems = (int64_t)counters[6];
    out->pairs_swept_one_sided = (int64_t)counters[7];
    out->candidates = (int64_t)win.verified;
    out->items = (int32_t)std::min<int64_t>(J.n_items, INT32_MAX);
    out->items_total = J.n_items;
    out->rows_per_item = J.rows_per_item;
    return DTWJ_OK;
}

// Candidate slots for a re-run that saw `need` pushes: a quarter more, at least twice the old.
int64_t regrown_cap(int64_t old_cap, int64_t need) {
    int64_t c = std::max<int64_t>(2 * old_cap, need + need / 4 + 1024);
    return std::min<int64_t>(c, kMaxCandCap);
}

// Copies the error model the set-up kernel computed into the host copy (used by finish_job).
int fetch_error_model(Job &J, const JoinArgs *d_args, hipStream_t s) {
    JoinArgs Ad;
    DTWJ_HIP(hipMemcpyAsync(&Ad, d_args, sizeof(Ad), hipMemcpyDeviceToHost, s));
    DTWJ_HIP(hipStreamSynchronize(s));
    J.A.amax = Ad.amax;
    J.A.eps_z = Ad.eps_z;
    J.A.rel_dtw = Ad.rel_dtw;
    J.A.valid = Ad.valid;
    return 0;
}

// Claim-order multiplier of a work list of `total` items: near the golden-ratio fraction of the
// list and coprime with its length (claim k -> item k * mul mod total is then a bijection that
// spreads the first claims over the whole list).  0 (index order) for lists of 2^32 items and
// more, where the 64-bit product could overflow.
int64_t claim_multiplier(int64_t total) {
    if (total < 3 || total >= ((int64_t)1 << 32)) return 0;
    int64_t mul = (int64_t)((double)total * 0.6180339887498949) | 1;
    while (std::gcd(mul, total) != 1) mul += 2;
    return mul % total;
}

int32_t grid_for(const dtwj_context *c, int32_t wpc, int64_t n_items) {
    wpc = wpc > 0 ? wpc : 8;
    int64_t grid = (int64_t)c->num_cus * wpc;
    if (grid > n_items) grid = n_items;
    return (int32_t)std::max<int64_t>(grid, 1);
}

int ensure_mem(dtwj_context *c, size_t need) {
    if (need > c->mem_cap && c->fixed) {
        char buf[200];
        std::snprintf(buf, sizeof(buf), "input too large for the fixed workspace: %zu bytes needed, %zu available",
                      need, c->mem_cap);
        return fail(DTWJ_EINVAL, buf);
    }
    if (need > c->mem_cap) {
        if (c->mem) DTWJ_HIP(hipFree(c->mem));
        c->mem = nullptr;
        c->mem_cap = 0;
        hipError_t e = hipMalloc(&c->mem, need);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(workspace)", __FILE__, __LINE__);
        c->mem_cap = need;
    }
    return 0;
}


// ---------------------------------------------------------------- learned proposal
// Device buffers of one proposal pass (candidates, raw features, network input and output).
struct Proposal {
    Job J;
    int2 *d_cand = nullptr;  // n_seed * 64 (scan output), then the de-duplicated list
    float *d_F = nullptr;    // count x DTWJ_MLP_IN raw features
    uint16_t *d_X = nullptr; // rows16 x pad BF16 network input
    float *d_P = nullptr;    // network parameters (FP32)
    uint16_t *d_W1b = nullptr;
    float *d_y = nullptr;    // network output
    float *d_key = nullptr;  // ranking keys: learned, then ED
    float *d_norm = nullptr; // feature mean, scale
    int64_t *d_vi = nullptr, *d_vj = nullptr;
    double *d_vcost = nullptr;
    uint64_t *d_keys = nullptr;  // candidate keys, sorted in place (pair_sort_len entries)
    std::vector<int2> cands;  // candidate pairs in (i, j) order (host)
};

// Statistics, candidate scan and raw features of one series (device-resident series d_a).
int propose(dtwj_context *c, const double *d_a, int64_t n, int32_t m, int32_t w, int32_t n_items, int32_t n_verify,
            Proposal &Pr) {
    Job &J = Pr.J;
    plan_job(J, d_a, n, m, w, 0, n_items > 0 ? n_items : 4096, true, {}, 1);
    const size_t ncand = (size_t)J.n_seed * 64;
    const size_t rows16 = (ncand + 15) / 16 * 16;
    const int np = mlp_param_count(), pad = mlp_feature_pad();
    const int nv = std::max(n_verify, 1) * 2;
    const int64_t sort_len = pair_sort_len((int64_t)ncand);
    auto carve = [&](Arena &ar) {
        take_job(ar, J);
        Pr.d_cand = ar.take<int2>(ncand + 1);
        Pr.d_F = ar.take<float>(ncand * DTWJ_MLP_IN + 1);
        Pr.d_X = ar.take<uint16_t>(rows16 * pad + 16);
        Pr.d_P = ar.take<float>(np);
        Pr.d_W1b = ar.take<uint16_t>((size_t)DTWJ_MLP_HID * pad);
        Pr.d_y = ar.take<float>(rows16 + 16);
        Pr.d_key = ar.take<float>(2 * (rows16 + 16));
        Pr.d_norm = ar.take<float>(2 * DTWJ_MLP_IN);
        Pr.d_vi = ar.take<int64_t>(nv);
        Pr.d_vj = ar.take<int64_t>(nv);
        Pr.d_vcost = ar.take<double>(nv);
        Pr.d_keys = ar.take<uint64_t>((size_t)sort_len);
    };
    Arena dry;
    carve(dry);
    if (int rc = ensure_mem(c, dry.off + 256)) return rc;
    Arena ar;
    ar.base = (char *)c->mem;
    ar.cap = c->mem_cap;
    carve(ar);
    hipStream_t s = c->stream;
    if (int rc = upload_job(J, s)) return rc;
    if (int rc = prep_series(J.P, s)) return rc;
    int32_t flags[4];
    DTWJ_HIP(hipMemcpyAsync(flags, J.P.flags, sizeof(flags), hipMemcpyDeviceToHost, s));
    DTWJ_HIP(hipStreamSynchronize(s));
    if (flags[0]) return fail(DTWJ_EINVAL, "series contains NaN or Inf");
    if (flags[1]) return fail(DTWJ_EINVAL, "a window of length m is constant (standard deviation 0)");
    SeedArgs S;
    S.a = J.d_a;
    S.mu = J.P.mu;
    S.win = J.P.win;
    S.n = J.n;
    S.nw = J.nw;
    S.m = J.m;
    S.item_start = J.d_item_start;
    S.groups = J.groups;
    S.n_items = J.n_items;
    S.rows_per_item = J.rows_per_item;
    S.rows = std::min(J.rows_per_item, 512);
    S.n_seed = J.n_seed;
    if (int rc = launch_seed_lanes(S, Pr.d_cand, s)) return rc;
    // the lanes' pairs in (i, j) order, sorted on the device; duplicates are dropped (none
    // arise: every (item, lane) of the scan owns its own rows and diagonals)
    if (int rc = launch_pair_sort(Pr.d_cand, (int64_t)ncand, Pr.d_keys, sort_len, s)) return rc;
    std::vector<uint64_t> keys(ncand);
    DTWJ_HIP(hipMemcpyAsync(keys.data(), Pr.d_keys, sizeof(uint64_t) * ncand, hipMemcpyDeviceToHost, s));
    DTWJ_HIP(hipStreamSynchronize(s));
    Pr.cands.clear();
    Pr.cands.reserve(ncand);
    for (size_t k = 0; k < ncand && keys[k] != ~0ull; ++k)
        if (k == 0 || keys[k] != keys[k - 1]) Pr.cands.push_back(make_int2((int)(keys[k] >> 32), (int)(uint32_t)keys[k]));
    const int cnt = (int)Pr.cands.size();
    if (cnt > 0) {
        DTWJ_HIP(hipMemcpyAsync(Pr.d_cand, Pr.cands.data(), sizeof(int2) * cnt, hipMemcpyHostToDevice, s));
        if (int rc = launch_prop_features(J.A, Pr.d_cand, cnt, Pr.d_F, s)) return rc;
    }
    return DTWJ_OK;
}

// Network output for the first `count` rows of d_F.
int score(dtwj_context *c, const dtwj_mlp &M, const float *d_F, int count, Proposal &Pr) {
    hipStream_t s = c->stream;
    const int rows16 = (count + 15) / 16 * 16;
    std::vector<float> P(mlp_param_count()), norm(2 * DTWJ_MLP_IN);
    mlp_to_params(M, P.data());
    for (int f = 0; f < DTWJ_MLP_IN; ++f) {
        norm[f] = M.feat_mean[f];
        norm[DTWJ_MLP_IN + f] = M.feat_scale[f];
    }
    DTWJ_HIP(hipMemcpyAsync(Pr.d_P, P.data(), sizeof(float) * P.size(), hipMemcpyHostToDevice, s));
    DTWJ_HIP(hipMemcpyAsync(Pr.d_norm, norm.data(), sizeof(float) * norm.size(), hipMemcpyHostToDevice, s));
    if (int rc = launch_params_to_bf16(Pr.d_P, Pr.d_W1b, s)) return rc;
    if (int rc = launch_prop_pack(d_F, count, rows16, Pr.d_norm, Pr.d_norm + DTWJ_MLP_IN, Pr.d_X, s)) return rc;
    return launch_mlp_forward(Pr.d_X, count, Pr.d_W1b, Pr.d_P, Pr.d_y, s);
}

// Exact (FP64) DTW cost of a pair list of the job's series.
int verify_pairs(dtwj_context *c, Job &J, const std::vector<int2> &pairs, Proposal &Pr, std::vector<double> &cost) {
    hipStream_t s = c->stream;
    const int64_t np = (int64_t)pairs.size();
    cost.assign(np, 0.0);
    if (np == 0) return DTWJ_OK;
    std::vector<int64_t> vi(np), vj(np);
    for (int64_t k = 0; k < np; ++k) {
        vi[k] = pairs[k].x;
        vj[k] = pairs[k].y;
    }
    DTWJ_HIP(hipMemcpyAsync(Pr.d_vi, vi.data(), sizeof(int64_t) * np, hipMemcpyHostToDevice, s));
    DTWJ_HIP(hipMemcpyAsync(Pr.d_vj, vj.data(), sizeof(int64_t) * np, hipMemcpyHostToDevice, s));
    if (int rc = launch_dtw_pairs(J.A, J.P.sig, Pr.d_vi, Pr.d_vj, np, 64, Pr.d_vcost, s)) return rc;
    DTWJ_HIP(hipMemcpyAsync(cost.data(), Pr.d_vcost, sizeof(double) * np, hipMemcpyDeviceToHost, s));
    DTWJ_HIP(hipStreamSynchronize(s));
    return DTWJ_OK;
}

// Device buffers of the network on their own (forward, gradient, training).
struct MlpBuffers {
    float *F = nullptr, *T = nullptr, *P = nullptr, *G = nullptr, *M1 = nullptr, *M2 = nullptr, *loss = nullptr;
    float *y = nullptr, *norm = nullptr;
    uint16_t *X = nullptr, *W1b = nullptr;
    int32_t *perm = nullptr;
};

int mlp_setup(dtwj_context *c, const dtwj_mlp &M, const float *feats, const float *targets, int64_t count,
              MlpBuffers &B) {
    const int np = mlp_param_count(), pad = mlp_feature_pad();
    const size_t rows16 = ((size_t)count + 15) / 16 * 16;
    auto carve = [&](Arena &ar) {
        B.F = ar.take<float>((size_t)count * DTWJ_MLP_IN + 1);
        B.T = ar.take<float>((size_t)count + 1);
        B.P = ar.take<float>(np);
        B.G = ar.take<float>(np + 1);
        B.M1 = ar.take<float>(np);
        B.M2 = ar.take<float>(np);
        B.loss = ar.take<float>(1);
        B.y = ar.take<float>(rows16 + 16);
        B.norm = ar.take<float>(2 * DTWJ_MLP_IN);
        B.X = ar.take<uint16_t>(rows16 * pad + 16);
        B.W1b = ar.take<uint16_t>((size_t)DTWJ_MLP_HID * pad);
        B.perm = ar.take<int32_t>((size_t)count + 1);
    };
    Arena dry;
    carve(dry);
    if (int rc = ensure_mem(c, dry.off + 256)) return rc;
    Arena ar;
    ar.base = (char *)c->mem;
    ar.cap = c->mem_cap;
    carve(ar);
    hipStream_t s = c->stream;
    std::vector<float> P(np), norm(2 * DTWJ_MLP_IN);
    mlp_to_params(M, P.data());
    for (int f = 0; f < DTWJ_MLP_IN; ++f) {
        norm[f] = M.feat_mean[f];
        norm[DTWJ_MLP_IN + f] = M.feat_scale[f];
    }
    DTWJ_HIP(hipMemcpyAsync(B.F, feats, sizeof(float) * count * DTWJ_MLP_IN, hipMemcpyHostToDevice, s));
    if (targets) DTWJ_HIP(hipMemcpyAsync(B.T, targets, sizeof(float) * count, hipMemcpyHostToDevice, s));
    DTWJ_HIP(hipMemcpyAsync(B.P, P.data(), sizeof(float) * np, hipMemcpyHostToDevice, s));
    DTWJ_HIP(hipMemcpyAsync(B.norm, norm.data(), sizeof(float) * norm.size(), hipMemcpyHostToDevice, s));
    if (int rc = launch_params_to_bf16(B.P, B.W1b, s)) return rc;
    return launch_prop_pack(B.F, (int)count, (int)rows16, B.norm, B.norm + DTWJ_MLP_IN, B.X, s);
}

int check_mlp_args(const dtwj_mlp *model, const float *feats, int64_t count) {
    if (!model || !feats) return fail(DTWJ_EINVAL, "null model or features");
    if (count <= 0 || count > (int64_t)1 << 28) return fail(DTWJ_EINVAL, "feature row count out of range");
    return DTWJ_OK;
}

struct Seeds {
    std::vector<int2> pairs;  // explicit seed pairs (mp-seeded entry point)
    bool gpu_scan = true;
};

int run_join(dtwj_context *c, const double *d_a, int64_t n, int32_t m, int32_t w_in, const dtwj_options *opt,
             const Seeds &seeds, dtwj_result *out, clk::time_point t_entry) {
    if (opt && (opt->sweep < 0 || opt->sweep > 2)) return fail(DTWJ_EINVAL, "options.sweep must be 0, 1 or 2");
    if (opt && opt->cand_cap < 0) return fail(DTWJ_EINVAL, "options.cand_cap must be >= 0");
    if (opt && (opt->paa < 0 || opt->paa > 2)) return fail(DTWJ_EINVAL, "options.paa must be 0, 1 or 2");
    Job J;
    plan_job(J, d_a, n, m, w_in, opt ? opt->rows_per_item : 0, opt ? opt->seed_items : 0, seeds.gpu_scan, seeds.pairs,
             (opt && opt->cand_cap > 0) ? (int64_t)opt->cand_cap : kCandCap);
    J.sweep_mode = opt ? opt->sweep : 0;
    J.paa_seg = paa_seg_for(J.m, J.w, opt ? opt->paa : 0);
    const int32_t wpc = (opt && opt->waves_per_cu > 0) ? opt->waves_per_cu : 8;
    const int32_t grid = grid_for(c, wpc, J.n_items);
    int32_t reruns = 0;
    double seed_dist = -1.0;
    for (;;) {
        JoinArgs *d_args = nullptr;
        int64_t *d_item_base = nullptr;
        unsigned long long *d_work = nullptr;
        uint32_t *d_queue = nullptr;
        double *d_cost64 = nullptr;
        Winner *d_part = nullptr, *d_win = nullptr;
        auto carve = [&](Arena &ar) {
            take_job(ar, J);
            d_args = ar.take<JoinArgs>(1);
            d_item_base = ar.take<int64_t>(3);
            d_work = ar.take<unsigned long long>(1);
            d_queue = ar.take<uint32_t>((size_t)grid * kJoinQueueCap);
            d_cost64 = ar.take<double>((size_t)J.cand_cap);
            d_part = ar.take<Winner>((size_t)rescore_partials());
            d_win = ar.take<Winner>(1);
        };
        Arena dry;
        carve(dry);
        if (int rc = ensure_mem(c, dry.off + 256)) return rc;
        Arena ar;
        ar.base = (char *)c->mem;
        ar.cap = c->mem_cap;
        carve(ar);

        hipStream_t s = c->stream;
        const auto t0 = clk::now();
        if (int rc = upload_job(J, s)) return rc;
        const int64_t item_base[3] = {0, J.n_items, claim_multiplier(J.n_items)};
        DTWJ_HIP(hipMemcpyAsync(d_item_base, item_base, sizeof(item_base), hipMemcpyHostToDevice, s));
        // ---- preprocessing
        if (int rc = prep_series(J.P, s)) return rc;
        int32_t flags[4];
        DTWJ_HIP(hipMemcpyAsync(flags, J.P.flags, sizeof(flags), hipMemcpyDeviceToHost, s));
        DTWJ_HIP(hipStreamSynchronize(s));
        if (flags[0]) return fail(DTWJ_EINVAL, "series contains NaN or Inf");
        if (flags[1]) return fail(DTWJ_EINVAL, "a window of length m is constant (standard deviation 0)");
        const auto t1 = clk::now();
        // ---- join descriptor, error model, resets; seed
        DTWJ_HIP(hipMemcpyAsync(d_args, &J.A, sizeof(JoinArgs), hipMemcpyHostToDevice, s));
        if (int rc = launch_join_setup(d_args, 1, d_work, s)) return rc;
        if (int rc = fetch_error_model(J, d_args, s)) return rc;
        if (int rc = enqueue_seeds(J, d_args, s)) return rc;
        if (seed_dist < 0) {
            // the seeds' best FP32 DTW cost, from their candidate records (the best-so-far word
            // holds a bound on that pair's exact cost, slightly above it)
            unsigned long long nseed = 0;
            DTWJ_HIP(hipMemcpyAsync(&nseed, J.d_ncands, sizeof(nseed), hipMemcpyDeviceToHost, s));
            DTWJ_HIP(hipStreamSynchronize(s));
            nseed = std::min<unsigned long long>(nseed, (unsigned long long)J.cand_cap);
            std::vector<int4> sc((size_t)nseed);
            if (nseed > 0) {
                DTWJ_HIP(hipMemcpyAsync(sc.data(), J.d_cands, nseed * sizeof(int4), hipMemcpyDeviceToHost, s));
                DTWJ_HIP(hipStreamSynchronize(s));
            }
            float best = 3.0e38f;
            for (const int4 &c : sc) best = std::min(best, __builtin_bit_cast(float, c.z));
            seed_dist = std::sqrt((double)best);
        }
        const auto t2 = clk::now();
        // ---- fused join
        if (c->ev0) DTWJ_HIP(hipEventRecord(c->ev0, s));
        if (int rc = launch_join(d_args, d_item_base, 1, J.m, J.w, d_work, d_queue, grid, s)) return rc;
        if (c->ev1) DTWJ_HIP(hipEventRecord(c->ev1, s));
        DTWJ_HIP(hipStreamSynchronize(s));
        const auto t3 = clk::now();
        float join_ms = (float)(secs(t2, t3) * 1e3);  // host clock without events
        if (c->ev0 && c->ev1) DTWJ_HIP(hipEventElapsedTime(&join_ms, c->ev0, c->ev1));
        // ---- FP64 re-verification of every candidate within the FP32 error of the best
        const int rc = finish_job(J, d_cost64, d_part, d_win, s, out);
        if (rc == kRegrow) {
            // more near-ties than slots: run again from this run's best-so-far (a valid
            // threshold, so the re-run pushes no more than this run did) with room for all
            if (c->fixed)
                return fail(DTWJ_EOVERFLOW, "more near-tie candidates than the fixed workspace holds "
                                            "(the allocating entry points regrow their buffer)");
            if (J.cand_cap >= kMaxCandCap || J.cand_need > kMaxCandCap)
                return fail(DTWJ_EOVERFLOW, "more than 2^31 - 1 near-tie candidates");
            J.cand_cap = regrown_cap(J.cand_cap, J.cand_need);
            ++reruns;
            continue;
        }
        if (rc != DTWJ_OK) return rc;
        const auto t4 = clk::now();
        out->seed_dist = seed_dist;
        out->t_prep = secs(t0, t1);
        out->t_seed = secs(t1, t2);
        out->t_join = secs(t2, t3);
        out->t_verify = secs(t3, t4);
        out->t_total = secs(t_entry, t4);
        out->join_kernel_ms = join_ms;
        out->waves_per_cu = wpc;
        out->reruns = reruns;
        if (opt && opt->verbose) {
            std::fprintf(stderr,
                         "[dtwj] n=%lld m=%d w=%d items=%lld swept=%lld ball=%lld kim=%lld keogh=%lld dtw_ab=%lld "
                         "cells=%lld cands=%lld seed=%.6f best=%.6f join=%.3f ms reruns=%d\n",
                         (long long)n, m, J.w, (long long)J.n_items, (long long)out->pairs_swept,
                         (long long)out->pass_ball, (long long)out->pass_kim, (long long)out->pass_keogh,
                         (long long)out->dtw_abandoned, (long long)out->dtw_cells, (long long)out->candidates,
                         out->seed_dist, out->dist, join_ms, reruns);
        }
        return DTWJ_OK;
    }
}

// ---------------------------------------------------------------- batched form
// Every series gets its own statistics, tables, best-so-far and candidate list; their work
// items are concatenated into one list (item_base) that one persistent join launch drains.
// Preprocessing, set-up, seeds and the join are enqueued on one stream, captured in one HIP
// graph unless opt->graph < 0.  Per-series validation failures are reported in status[] and
// the series is left out; *join_ms receives the device time of the captured work.
int run_batch(dtwj_context *c, const dtwj_series *series, int32_t count, const dtwj_options *opt, int32_t rows_req,
              int32_t wpc, dtwj_result *out, int32_t *status, float *join_ms) {
    if (opt && (opt->sweep < 0 || opt->sweep > 2)) return fail(DTWJ_EINVAL, "options.sweep must be 0, 1 or 2");
    if (opt && (opt->paa < 0 || opt->paa > 2)) return fail(DTWJ_EINVAL, "options.paa must be 0, 1 or 2");
    std::vector<Job> jobs;
    std::vector<int32_t> owner;  // job -> batch index
    jobs.reserve(count);
    for (int32_t k = 0; k < count; ++k) {
        std::memset(&out[k], 0, sizeof(dtwj_result));
        const dtwj_series &S = series[k];
        status[k] = validate(S.data, S.n, S.m, S.w);
        if (status[k] == DTWJ_OK) {
            for (int64_t t = 0; t < S.n; ++t)
                if (!std::isfinite(S.data[t])) {
                    status[k] = fail(DTWJ_EINVAL, "series contains NaN or Inf");
                    break;
                }
        }
        if (status[k] != DTWJ_OK) continue;
        jobs.emplace_back();
        // batches keep 4096 seed items per series: their seeds run in one launch, and the
        // shorter series of a batch gain more from a close seed (measured: 11.6 s vs 11.9 s)
        const int32_t seed_req = opt && opt->seed_items > 0 ? opt->seed_items : 4096;
        plan_job(jobs.back(), nullptr, S.n, S.m, S.w, rows_req, seed_req, true, {}, kBatchCandCap);
        jobs.back().sweep_mode = opt ? opt->sweep : 0;
        jobs.back().paa_seg = paa_seg_for(jobs.back().m, jobs.back().w, opt ? opt->paa : 0);
        owner.push_back(k);
    }
    const int32_t nj = (int32_t)jobs.size();
    if (nj == 0) return DTWJ_OK;
    std::vector<int64_t> item_base(nj + 1, 0);
    int32_t max_m = 0, max_w = 0;
    for (int32_t t = 0; t < nj; ++t) {
        item_base[t + 1] = item_base[t] + jobs[t].n_items;
        max_m = std::max(max_m, jobs[t].m);
        max_w = std::max(max_w, jobs[t].w);
    }
    const int32_t grid = grid_for(c, wpc, item_base[nj]);
    std::vector<double *> d_data(nj);
    JoinArgs *d_args = nullptr;
    int64_t *d_item_base = nullptr;
    unsigned long long *d_work = nullptr;
    uint32_t *d_queue = nullptr;
    double *d_cost64 = nullptr;
    Winner *d_part = nullptr, *d_win = nullptr;
    SeedJob *d_sjobs = nullptr;
    SeedDtwJob *d_djobs = nullptr;
    int32_t *d_sbase = nullptr, *d_dbase = nullptr;
    auto carve = [&](Arena &ar) {
        for (int32_t t = 0; t < nj; ++t) {
            d_data[t] = ar.take<double>(jobs[t].n);
            jobs[t].d_a = d_data[t];
            take_job(ar, jobs[t]);
        }
        d_args = ar.take<JoinArgs>(nj);
        d_item_base = ar.take<int64_t>(nj + 2);
        d_work = ar.take<unsigned long long>(1);
        d_queue = ar.take<uint32_t>((size_t)grid * kJoinQueueCap);
        d_cost64 = ar.take<double>(kBatchCandCap);
        d_part = ar.take<Winner>((size_t)rescore_partials());
        d_win = ar.take<Winner>(1);
        d_sjobs = ar.take<SeedJob>(nj);
        d_sbase = ar.take<int32_t>(nj + 1);
        d_djobs = ar.take<SeedDtwJob>(nj);
        d_dbase = ar.take<int32_t>(nj + 4);
    };
    Arena dry;
    carve(dry);
    if (int rc = ensure_mem(c, dry.off + 256)) return rc;
    Arena ar;
    ar.base = (char *)c->mem;
    ar.cap = c->mem_cap;
    carve(ar);

    // Seeds of all series in few launches: one seed scan over every series' blocks, and one
    // seed-pair DTW launch per band class (offset pairs per lane), each block finding its
    // series from prefix sums of the blocks.
    std::vector<SeedJob> hsj;
    std::vector<int32_t> hsb{0};
    for (int32_t t = 0; t < nj; ++t) {
        if (jobs[t].n_seed <= 0) continue;
        SeedJob q;
        q.S = seed_args(jobs[t]);
        q.pairs = jobs[t].d_seed_pairs;
        hsj.push_back(q);
        hsb.push_back(hsb.back() + jobs[t].n_seed + 1);
    }
    struct DtwClass {
        int32_t job0, cnt, base0, blocks, max_m, w;
    };
    std::vector<DtwClass> dcls;
    std::vector<SeedDtwJob> hdj;
    std::vector<int32_t> hdb;
    auto qp_of = [](int32_t w) {  // offset pairs per lane of the DTW kernels: 128 QP >= 2w + 1
        const int32_t band = 2 * w + 1;
        return band <= 128 ? 1 : band <= 256 ? 2 : band <= 512 ? 4 : 8;
    };
    for (int qp : {1, 2, 4, 8}) {
        DtwClass k{(int32_t)hdj.size(), 0, (int32_t)hdb.size(), 0, 0, 0};
        hdb.push_back(0);
        for (int32_t t = 0; t < nj; ++t) {
            if (qp_of(jobs[t].w) != qp || jobs[t].n_seed_pairs <= 0) continue;
            hdj.push_back(SeedDtwJob{jobs[t].d_seed_pairs, t, jobs[t].n_seed_pairs});
            k.blocks += jobs[t].n_seed_pairs;
            hdb.push_back(k.blocks);
            k.max_m = std::max(k.max_m, jobs[t].m);
            k.w = jobs[t].w;
            ++k.cnt;
        }
        if (k.cnt > 0)
            dcls.push_back(k);
        else
            hdb.pop_back();
    }

    hipStream_t s = c->stream;
    if (!hsj.empty()) {
        DTWJ_HIP(hipMemcpyAsync(d_sjobs, hsj.data(), sizeof(SeedJob) * hsj.size(), hipMemcpyHostToDevice, s));
        DTWJ_HIP(hipMemcpyAsync(d_sbase, hsb.data(), sizeof(int32_t) * hsb.size(), hipMemcpyHostToDevice, s));
    }
    if (!hdj.empty()) {
        DTWJ_HIP(hipMemcpyAsync(d_djobs, hdj.data(), sizeof(SeedDtwJob) * hdj.size(), hipMemcpyHostToDevice, s));
        DTWJ_HIP(hipMemcpyAsync(d_dbase, hdb.data(), sizeof(int32_t) * hdb.size(), hipMemcpyHostToDevice, s));
    }
    std::vector<JoinArgs> hargs(nj);
    for (int32_t t = 0; t < nj; ++t) {
        const dtwj_series &S = series[owner[t]];
        DTWJ_HIP(hipMemcpyAsync(d_data[t], S.data, sizeof(double) * S.n, hipMemcpyHostToDevice, s));
        if (int rc = upload_job(jobs[t], s)) return rc;
        hargs[t] = jobs[t].A;
    }
    DTWJ_HIP(hipMemcpyAsync(d_args, hargs.data(), sizeof(JoinArgs) * nj, hipMemcpyHostToDevice, s));
    item_base.push_back(claim_multiplier(item_base[nj]));
    DTWJ_HIP(hipMemcpyAsync(d_item_base, item_base.data(), sizeof(int64_t) * (nj + 2), hipMemcpyHostToDevice, s));
    DTWJ_HIP(hipStreamSynchronize(s));

    auto enqueue_all = [&]() -> int {
        for (int32_t t = 0; t < nj; ++t)
            if (int rc = prep_series(jobs[t].P, s)) return rc;
        if (int rc = launch_join_setup(d_args, nj, d_work, s)) return rc;
        if (int rc = launch_seed_batch(d_sjobs, d_sbase, (int32_t)hsj.size(), hsb.back(), s)) return rc;
        for (const DtwClass &k : dcls)
            if (int rc = launch_seed_dtw_batch(d_args, k.max_m, k.w, d_djobs + k.job0, d_dbase + k.base0, k.cnt,
                                               k.blocks, s))
                return rc;
        return launch_join(d_args, d_item_base, nj, max_m, max_w, d_work, d_queue, grid, s);
    };
    const bool use_graph = !(opt && opt->graph < 0);
    DTWJ_HIP(hipEventRecord(c->ev0, s));
    if (use_graph) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        DTWJ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_all();
        hipError_t ec = hipStreamEndCapture(s, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (ec != hipSuccess) return hip_fail(ec, "hipStreamEndCapture", __FILE__, __LINE__);
        hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return hip_fail(ei, "hipGraphInstantiate", __FILE__, __LINE__);
        }
        DTWJ_HIP(hipEventRecord(c->ev0, s));
        hipError_t el = hipGraphLaunch(exec, s);
        DTWJ_HIP(hipEventRecord(c->ev1, s));
        hipError_t es = hipStreamSynchronize(s);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        if (el != hipSuccess) return hip_fail(el, "hipGraphLaunch", __FILE__, __LINE__);
        if (es != hipSuccess) return hip_fail(es, "batch graph", __FILE__, __LINE__);
    } else {
        if (int rc = enqueue_all()) return rc;
        DTWJ_HIP(hipEventRecord(c->ev1, s));
        DTWJ_HIP(hipStreamSynchronize(s));
    }
    float ms = 0.f;
    DTWJ_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (join_ms) *join_ms = ms;
    std::vector<int32_t> regrow;  // batch indices whose candidates overflowed their buffer
    for (int32_t t = 0; t < nj; ++t) {
        const int32_t k = owner[t];
        int32_t flags[4];
        DTWJ_HIP(hipMemcpyAsync(flags, jobs[t].P.flags, sizeof(flags), hipMemcpyDeviceToHost, s));
        if (int rc = fetch_error_model(jobs[t], d_args + t, s)) return rc;
        if (flags[0] || flags[1]) {
            status[k] = fail(DTWJ_EINVAL, flags[0] ? "series contains NaN or Inf"
                                                   : "a window of length m is constant (standard deviation 0)");
            continue;
        }
        status[k] = finish_job(jobs[t], d_cost64, d_part, d_win, s, &out[k]);
        if (status[k] == kRegrow) {
            regrow.push_back(k);
            continue;
        }
        out[k].join_kernel_ms = ms;
        out[k].waves_per_cu = wpc;
    }
    // A series with more near-ties than its batch buffer is joined again on its own (the
    // single-series path regrows its buffer as needed); the batch's device data is no longer
    // used, so the context's workspace can be reused.
    for (int32_t k : regrow) {
        const dtwj_series &S = series[k];
        const double *d = nullptr;
        int rc = dtwj_context_upload(c, S.data, S.n, &d);
        if (rc == DTWJ_OK) {
            dtwj_options o = opt ? *opt : dtwj_options{};
            o.rows_per_item = rows_req;
            o.waves_per_cu = wpc;
            rc = run_join(c, d, S.n, S.m, S.w, &o, Seeds{}, &out[k], clk::now());
            if (rc == DTWJ_OK) out[k].reruns += 1;
        }
        status[k] = rc;
    }
    return DTWJ_OK;
}

// Autotuner of the batched form: times the join on a sample of the batch (up to 32 series
// spread over the batch, in its order) for each candidate tile height and occupancy, and
// keeps the fastest.  The choice is cached in the context, keyed by the batch's (m, w) mix
// and size, so repeated batches of the same shape do not pay for it again.
int autotune_batch(dtwj_context *c, const dtwj_series *series, int32_t count, const dtwj_options *opt,
                   int32_t *rows_out, int32_t *wpc_out) {
    std::vector<std::pair<int32_t, int32_t>> mw;
    int64_t total = 0;
    for (int32_t k = 0; k < count; ++k) {
        mw.emplace_back(series[k].m, series[k].w);
        total += series[k].n;
    }
    std::sort(mw.begin(), mw.end());
    mw.erase(std::unique(mw.begin(), mw.end()), mw.end());
    std::string key = std::to_string(count) + ":" + std::to_string(total >> 16);
    for (auto &p : mw) key += ":" + std::to_string(p.first) + "/" + std::to_string(p.second);
    auto it = c->tuned.find(key);
    if (it != c->tuned.end()) {
        *rows_out = it->second.first;
        *wpc_out = it->second.second;
        return DTWJ_OK;
    }
    // Every candidate is timed on the whole batch: a part of it runs fewer items per wave and
    // so a different tail of the persistent launch, and misleads (on the 512-series workload a
    // 32-series sample picked 4096-row items and a 128-series one 1024-row items, 3.7 % and
    // 3.2 % slower on the whole batch than the default 2048).  A candidate replaces the default
    // (first) only when it is 3 % faster: on that workload the candidates lie within 3 % of
    // each other, about the spread between two runs of one choice, so the default stays.
    // Tuning costs one join per candidate, once per context and batch shape.
    const std::vector<dtwj_series> sample(series, series + count);
    const int32_t ns = count;
    const int32_t cand[][2] = {{2048, 8}, {1024, 8}, {3072, 8}, {2048, 6}};
    std::vector<dtwj_result> res(ns);
    std::vector<int32_t> st(ns);
    dtwj_options o = opt ? *opt : dtwj_options{};
    o.graph = -1;
    float best = 0.f;
    int32_t br = 2048, bw = 8;
    bool have = false;
    for (const auto &cf : cand) {
        float ms = 0.f;
        if (int rc = run_batch(c, sample.data(), ns, &o, cf[0], cf[1], res.data(), st.data(), &ms)) return rc;
        if (!have || ms < best * (br == cand[0][0] && bw == cand[0][1] ? 0.97f : 1.0f)) {
            best = ms;
            br = cf[0];
            bw = cf[1];
            have = true;
        }
    }
    c->tuned[key] = {br, bw};
    *rows_out = br;
    *wpc_out = bw;
    return DTWJ_OK;
}

int with_context(int32_t device, dtwj_context **c) { return dtwj_context_create(device, c); }

}  // namespace

extern "C" {

int dtwj_plan(int64_t n, int32_t m, int32_t w, const dtwj_options *opt, dtwj_plan_info *out) {
    if (!out) return fail(DTWJ_EINVAL, "null plan");
    std::memset(out, 0, sizeof(*out));
    // validate() reads no sample: a non-null placeholder stands for the series
    static const double placeholder = 0.0;
    if (int rc = validate(&placeholder, n, m, w)) return rc;
    Job J;
    plan_job(J, nullptr, n, m, w, opt ? opt->rows_per_item : 0, opt ? opt->seed_items : 0, true, {}, kCandCap);
    if (opt && (opt->paa < 0 || opt->paa > 2)) return fail(DTWJ_EINVAL, "options.paa must be 0, 1 or 2");
    J.paa_seg = paa_seg_for(J.m, J.w, opt ? opt->paa : 0);
    const int32_t wpc = (opt && opt->waves_per_cu > 0) ? opt->waves_per_cu : 8;
    const int64_t grid = std::min<int64_t>((int64_t)256 * wpc, std::max<int64_t>(J.n_items, 1));
    auto pad = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t bytes = job_bytes(J);
    bytes += pad(sizeof(JoinArgs)) + pad(24) + pad(8) + pad(4 * (size_t)grid * kJoinQueueCap) +
             pad(8 * (size_t)J.cand_cap) + pad(sizeof(Winner) * (size_t)rescore_partials()) + pad(sizeof(Winner));
    out->pairs_total = J.K * (J.K + 1) / 2;
    out->items = J.n_items;
    out->rows_per_item = J.rows_per_item;
    out->diag_groups = J.groups;
    out->device_bytes = (int64_t)bytes;
    return DTWJ_OK;
}

int dtwj_top1_device_noalloc(const double *d_series, int64_t n, int32_t m, int32_t w, const dtwj_options *opt,
                             void *stream, dtwj_result *out) {
    const auto te = clk::now();
    if (!out) return fail(DTWJ_EINVAL, "null result");
    if (int rc = validate(d_series, n, m, w)) return rc;
    if (opt && (opt->cand_cap < 0 || opt->waves_per_cu < 0)) return fail(DTWJ_EINVAL, "bad options");
    const int32_t device = opt ? opt->device : 0;
    const int count = dtwj_device_count();
    if (count <= 0) return fail(DTWJ_EDEVICE, "no HIP device available");
    if (device < 0 || device >= count) return fail(DTWJ_EINVAL, "device ordinal out of range");
    DTWJ_HIP(hipSetDevice(device));
    void *ws = nullptr;
    if (int rc = noalloc_scratch(device, &ws)) return rc;
    int cus = 0;
    DTWJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    // a context over the fixed workspace and the caller's stream: nothing is created or freed
    dtwj_context c;
    c.device = device;
    c.stream = (hipStream_t)stream;
    c.mem = ws;
    c.mem_cap = noalloc_scratch_bytes();
    c.num_cus = cus;
    c.fixed = true;
    dtwj_options o;
    if (opt)
        o = *opt;
    else
        dtwj_default_options(&o);
    if (o.waves_per_cu == 0) o.waves_per_cu = 4;  // 1024 survivor queues of 16 KB on 256 CUs
    if (o.cand_cap == 0) o.cand_cap = 1 << 16;
    if (o.paa == 0) o.paa = 2;  // the segment-mean table (192 bytes per window) stays out of the workspace
    const int rc = run_join(&c, d_series, n, m, w, &o, Seeds{}, out, te);
    c.mem = nullptr;  // the workspace stays with the device
    return rc;
}

int dtwj_top1_device(dtwj_context *c, const double *d_series, int64_t n, int32_t m, int32_t w,
                     const dtwj_options *opt, dtwj_result *out) {
    const auto te = clk::now();
    if (!c || !out) return fail(DTWJ_EINVAL, "null context or result");
    if (int rc = validate(d_series, n, m, w)) return rc;
    DTWJ_HIP(hipSetDevice(c->device));
    Seeds seeds;
    return run_join(c, d_series, n, m, w, opt, seeds, out, te);
}

static int top1_host(const double *series, int64_t n, int32_t m, int32_t w, const Seeds &seeds,
                     const dtwj_options *opt, dtwj_result *out) {
    const auto te = clk::now();
    if (!out) return fail(DTWJ_EINVAL, "null result");
    if (int rc = validate(series, n, m, w)) return rc;
    for (int64_t t = 0; t < n; ++t)
        if (!std::isfinite(series[t])) return fail(DTWJ_EINVAL, "series contains NaN or Inf");
    dtwj_context *c = nullptr;
    if (int rc = with_context(opt ? opt->device : 0, &c)) return rc;
    const double *d = nullptr;
    int rc = dtwj_context_upload(c, series, n, &d);
    if (rc == DTWJ_OK) rc = run_join(c, d, n, m, w, opt, seeds, out, te);
    dtwj_context_destroy(c);
    return rc;
}

int dtwj_top1(const double *series, int64_t n, int32_t m, int32_t w, const dtwj_options *opt, dtwj_result *out) {
    Seeds seeds;
    return top1_host(series, n, m, w, seeds, opt, out);
}

int dtwj_top1_seeded(const double *series, int64_t n, int32_t m, int32_t w, const double *mp_ed,
                     const int32_t *mp_idx_1based, int64_t n_mp, const dtwj_options *opt, dtwj_result *out) {
    if (!mp_ed || !mp_idx_1based) return fail(DTWJ_EINVAL, "mp_ed / mp_idx is NULL");
    if (n_mp != n - m + 1) return fail(DTWJ_EINVAL, "matrix profile length must be n - m + 1");
    // the 100 smallest profile entries (reference k = 100) in scan order
    const int k = 100;
    std::vector<double> sorted(mp_ed, mp_ed + n_mp);
    std::vector<double> finite;
    for (double x : sorted)
        if (std::isfinite(x)) finite.push_back(x);
    Seeds seeds;
    seeds.gpu_scan = false;
    if (!finite.empty()) {
        const size_t kk = std::min<size_t>(k, finite.size());
        std::nth_element(finite.begin(), finite.begin() + (kk - 1), finite.end());
        const double kth = finite[kk - 1];
        for (int64_t t = 0; t < n_mp && (int)seeds.pairs.size() < k; ++t) {
            if (!(mp_ed[t] <= kth)) continue;
            int64_t a = t, b = (int64_t)mp_idx_1based[t] - 1;
            if (b < 0 || b >= n_mp) continue;
            if (a > b) std::swap(a, b);
            if (b - a < m) continue;
            seeds.pairs.push_back(make_int2((int)a, (int)b));
        }
    }
    return top1_host(series, n, m, w, seeds, opt, out);
}

// ED matrix profile of a series already on the device, into host arrays mp / mpi (nw each).
static int matrix_profile_on(dtwj_context *c, const double *d_a, int64_t n, int32_t m, double *mp, int32_t *mpi) {
    const int64_t nw = n - m + 1;
    const int64_t nsc = mp_scratch_doubles(n, m);
    double *d_scr = nullptr, *d_mp = nullptr;
    int32_t *d_mpi = nullptr, *d_flags = nullptr;
    auto carve = [&](Arena &ar) {
        d_scr = ar.take<double>((size_t)nsc);
        d_mp = ar.take<double>((size_t)nw);
        d_mpi = ar.take<int32_t>((size_t)nw);
        d_flags = ar.take<int32_t>(2);
    };
    Arena dry;
    carve(dry);
    if (int rc = ensure_mem(c, dry.off + 256)) return rc;
    Arena ar;
    ar.base = (char *)c->mem;
    ar.cap = c->mem_cap;
    carve(ar);
    hipStream_t s = c->stream;
    if (int rc = launch_matrix_profile(d_a, n, m, d_scr, d_flags, d_mp, d_mpi, s)) return rc;
    int32_t flags[2] = {0, 0};
    DTWJ_HIP(hipMemcpyAsync(flags, d_flags, sizeof(flags), hipMemcpyDeviceToHost, s));
    DTWJ_HIP(hipStreamSynchronize(s));
    if (flags[0]) return fail(DTWJ_EINVAL, "series contains NaN or Inf");
    if (flags[1]) return fail(DTWJ_EINVAL, "a window of length m is constant (standard deviation 0)");
    DTWJ_HIP(hipMemcpyAsync(mp, d_mp, sizeof(double) * nw, hipMemcpyDeviceToHost, s));
    DTWJ_HIP(hipMemcpyAsync(mpi, d_mpi, sizeof(int32_t) * nw, hipMemcpyDeviceToHost, s));
    DTWJ_HIP(hipStreamSynchronize(s));
    return DTWJ_OK;
}

// The reference's seed windows: the indices whose profile value is strictly below the k-th
// smallest value, in scan order, padded with index 0 to k + 10 entries, of which the seed loop
// visits the first k (min_v_k, GPU_version1/allmatrixop/min_v.cu:15-46, and the loop at
// GPU_version1/src/new_dtw_motifGUI_malloc.cu:82-107).
static std::vector<int64_t> reference_seed_windows(const double *mp, int64_t nw, int k) {
    std::vector<double> heap(k, std::numeric_limits<double>::infinity());  // max-heap of the k smallest
    for (int64_t t = 0; t < nw; ++t) {
        if (mp[t] < heap.front()) {
            std::pop_heap(heap.begin(), heap.end());
            heap.back() = mp[t];
            std::push_heap(heap.begin(), heap.end());
        }
    }
    const double kth = heap.front();
    std::vector<int64_t> out((size_t)k + 10, 0);
    size_t idx = 0;
    for (int64_t t = 0; t < nw && idx < out.size(); ++t)
        if (mp[t] < kth) out[idx++] = t;
    out.resize((size_t)k);
    return out;
}

int dtwj_matrix_profile(const double *series, int64_t n, int32_t m, double *mp_ed, int32_t *mp_idx_1based) {
    if (!mp_ed || !mp_idx_1based) return fail(DTWJ_EINVAL, "null output");
    if (int rc = validate(series, n, m, 0)) return rc;
    for (int64_t t = 0; t < n; ++t)
        if (!std::isfinite(series[t])) return fail(DTWJ_EINVAL, "series contains NaN or Inf");
    dtwj_context *c = nullptr;
    if (int rc = with_context(0, &c)) return rc;
    const double *d = nullptr;
    int rc = dtwj_context_upload(c, series, n, &d);
    if (rc == DTWJ_OK) rc = matrix_profile_on(c, d, n, m, mp_ed, mp_idx_1based);
    dtwj_context_destroy(c);
    return rc;
}

int dtwj_motif_discovery(const double *series, int64_t n, int32_t m, int32_t w, const char *time_file,
                         const dtwj_options *opt, dtwj_result *out) {
    dtwj_result local;
    dtwj_result *r = out ? out : &local;
    if (int rc = validate(series, n, m, w)) return rc;
    // ED matrix profile (mpx_v2 with minlag = m, GPU_version1/src/DTWMotifDiscoveryGUI.cu:20-24)
    const int64_t nw = n - m + 1;
    std::vector<double> mp((size_t)nw);
    std::vector<int32_t> mpi((size_t)nw);
    if (int rc = dtwj_matrix_profile(series, n, m, mp.data(), mpi.data())) return rc;
    // the reference's seed pairs: window t of the 100 smallest entries and its nearest neighbour
    const int k = 100;
    const std::vector<int64_t> win = reference_seed_windows(mp.data(), nw, k);
    std::vector<int64_t> pi, pj;
    for (int64_t t : win) {
        const int64_t p = (int64_t)mpi[(size_t)t] - 1;
        if (p < 0 || p >= nw) continue;
        pi.push_back(std::min(t, p));
        pj.push_back(std::max(t, p));
    }
    // FP64 DTW of every seed pair: the seed is min(mp_ed of the first seed window, their DTW)
    // (GPU_version1/src/new_dtw_motifGUI_malloc.cu:85-106)
    double seed = mp[(size_t)win[0]];
    if (!pi.empty()) {
        std::vector<double> cost(pi.size());
        if (int rc = dtwj_dtw_pairs(series, n, m, w, pi.data(), pj.data(), (int64_t)pi.size(), 64, cost.data()))
            return rc;
        for (double cst : cost) seed = std::min(seed, std::sqrt(cst));
    }
    Seeds seeds;
    seeds.gpu_scan = false;
    for (size_t q = 0; q < pi.size(); ++q) seeds.pairs.push_back(make_int2((int)pi[q], (int)pj[q]));
    if (int rc = top1_host(series, n, m, w, seeds, opt, r)) return rc;
    r->seed_dist = seed;
    // reference result lines (CPU_version1/new_dtw_motifGUI.cpp:575-579,
    // GPU_version1/src/new_dtw_motifGUI_malloc.cu:744-748); indices 1-based there
    std::printf("dtw cnt = %lld\n", (long long)(r->pass_keogh));
    std::printf("first min is %lld sec min is %lld\n", (long long)(r->i + 1), (long long)(r->j + 1));
    std::printf("total time is %fs   input of bsf:%5.3f  ,final distance :%f  \n", r->t_total, r->seed_dist, r->dist);
    std::printf("pair (0-based) i=%lld j=%lld dist=%.12g\n", (long long)r->i, (long long)r->j, r->dist);
    std::fflush(stdout);
    if (time_file) {
        FILE *op = std::fopen(time_file, "a");
        if (!op) return fail(DTWJ_EIO, std::string("cannot open time file: ") + time_file);
        std::fprintf(op, "%lf\n", r->t_total);
        std::fclose(op);
    }
    return DTWJ_OK;
}

int dtwj_context_top1_batch(dtwj_context *c, const dtwj_series *series, int32_t count, const dtwj_options *opt,
                            dtwj_result *out, int32_t *status) {
    const auto te = clk::now();
    if (!c || !series || count <= 0 || !out || !status) return fail(DTWJ_EINVAL, "bad batch arguments");
    DTWJ_HIP(hipSetDevice(c->device));
    int32_t rows = opt ? opt->rows_per_item : 0, wpc = opt ? opt->waves_per_cu : 0;
    if (opt && opt->autotune > 0)
        if (int rc = autotune_batch(c, series, count, opt, &rows, &wpc)) return rc;
    float ms = 0.f;
    if (int rc = run_batch(c, series, count, opt, rows, wpc > 0 ? wpc : 8, out, status, &ms)) return rc;
    const double total = secs(te, clk::now());
    for (int32_t k = 0; k < count; ++k) out[k].t_total = total;
    return DTWJ_OK;
}

int dtwj_top1_batch(const dtwj_series *series, int32_t count, const dtwj_options *opt, dtwj_result *out,
                    int32_t *status) {
    if (!series || count <= 0 || !out || !status) return fail(DTWJ_EINVAL, "bad batch arguments");
    dtwj_context *c = nullptr;
    if (int rc = with_context(opt ? opt->device : 0, &c)) return rc;
    int rc = dtwj_context_top1_batch(c, series, count, opt, out, status);
    dtwj_context_destroy(c);
    return rc;
}


int dtwj_proposal_candidates(const double *series, int64_t n, int32_t m, int32_t w, int32_t n_items, int64_t cap,
                             int64_t *pi, int64_t *pj, float *features, int64_t *count) {
    if (!pi || !pj || !features || !count || cap <= 0) return fail(DTWJ_EINVAL, "bad output arguments");
    if (int rc = validate(series, n, m, w)) return rc;
    dtwj_context *c = nullptr;
    if (int rc = with_context(0, &c)) return rc;
    const double *d = nullptr;
    Proposal Pr;
    int rc = dtwj_context_upload(c, series, n, &d);
    if (rc == DTWJ_OK) rc = propose(c, d, n, m, w, n_items, 1, Pr);
    if (rc == DTWJ_OK) {
        const int64_t cnt = std::min<int64_t>((int64_t)Pr.cands.size(), cap);
        for (int64_t k = 0; k < cnt; ++k) {
            pi[k] = Pr.cands[k].x;
            pj[k] = Pr.cands[k].y;
        }
        if (cnt > 0) {
            // the features were computed on the context's (non-blocking) stream
            hipError_t e = hipMemcpyAsync(features, Pr.d_F, sizeof(float) * cnt * DTWJ_MLP_IN, hipMemcpyDeviceToHost,
                                          c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = hip_fail(e, "features", __FILE__, __LINE__);
        }
        *count = cnt;
    }
    dtwj_context_destroy(c);
    return rc;
}

int dtwj_mlp_forward(const dtwj_mlp *model, const float *features, int64_t count, float *out) {
    if (int rc = check_mlp_args(model, features, count)) return rc;
    if (!out) return fail(DTWJ_EINVAL, "null output");
    dtwj_context *c = nullptr;
    if (int rc = with_context(0, &c)) return rc;
    MlpBuffers B;
    int rc = mlp_setup(c, *model, features, nullptr, count, B);
    if (rc == DTWJ_OK) rc = launch_mlp_forward(B.X, (int)count, B.W1b, B.P, B.y, c->stream);
    if (rc == DTWJ_OK) {
        hipError_t e = hipMemcpyAsync(out, B.y, sizeof(float) * count, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "mlp forward", __FILE__, __LINE__);
    }
    dtwj_context_destroy(c);
    return rc;
}

int dtwj_mlp_gradient(const dtwj_mlp *model, const float *features, const float *targets, int64_t count,
                      dtwj_mlp *grad, double *loss) {
    if (int rc = check_mlp_args(model, features, count)) return rc;
    if (!targets || !grad) return fail(DTWJ_EINVAL, "null targets or gradient");
    dtwj_context *c = nullptr;
    if (int rc = with_context(0, &c)) return rc;
    MlpBuffers B;
    int rc = mlp_setup(c, *model, features, targets, count, B);
    if (rc == DTWJ_OK) rc = launch_mlp_grad(B.X, B.T, nullptr, (int)count, B.W1b, B.P, B.G, B.loss, c->stream);
    if (rc == DTWJ_OK) {
        std::vector<float> G(mlp_param_count());
        float l = 0.f;
        hipError_t e = hipMemcpyAsync(G.data(), B.G, sizeof(float) * G.size(), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&l, B.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            rc = hip_fail(e, "mlp gradient", __FILE__, __LINE__);
        } else {
            std::memset(grad, 0, sizeof(*grad));
            params_to_mlp(G.data(), *grad);
            if (loss) *loss = l;
        }
    }
    dtwj_context_destroy(c);
    return rc;
}

int dtwj_mlp_train(dtwj_mlp *model, const float *features, const float *targets, int64_t count, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, double *final_loss) {
    if (int rc = check_mlp_args(model, features, count)) return rc;
    if (!targets || epochs <= 0 || batch <= 0 || !(lr > 0.f)) return fail(DTWJ_EINVAL, "bad training arguments");
    dtwj_context *c = nullptr;
    if (int rc = with_context(0, &c)) return rc;
    hipStream_t s = c->stream;
    MlpBuffers B;
    int rc = mlp_setup(c, *model, features, targets, count, B);
    const int np = mlp_param_count();
    std::vector<int32_t> perm((size_t)count);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(seed);
    int step = 0;
    if (rc == DTWJ_OK) {
        hipError_t e = hipMemsetAsync(B.M1, 0, sizeof(float) * np, s);
        if (e == hipSuccess) e = hipMemsetAsync(B.M2, 0, sizeof(float) * np, s);
        if (e != hipSuccess) rc = hip_fail(e, "adam state", __FILE__, __LINE__);
    }
    for (int32_t ep = 0; ep < epochs && rc == DTWJ_OK; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        hipError_t e = hipMemcpyAsync(B.perm, perm.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            rc = hip_fail(e, "permutation", __FILE__, __LINE__);
            break;
        }
        for (int64_t b0 = 0; b0 < count && rc == DTWJ_OK; b0 += batch) {
            const int bsz = (int)std::min<int64_t>(batch, count - b0);
            rc = launch_mlp_grad(B.X, B.T, B.perm + b0, bsz, B.W1b, B.P, B.G, B.loss, s);
            if (rc == DTWJ_OK) rc = launch_mlp_adam(B.P, B.G, B.M1, B.M2, lr, ++step, B.W1b, s);
        }
        // the host copy of the permutation is reused: wait before the next shuffle
        if (rc == DTWJ_OK) {
            e = hipStreamSynchronize(s);
            if (e != hipSuccess) rc = hip_fail(e, "training epoch", __FILE__, __LINE__);
        }
    }
    if (rc == DTWJ_OK) rc = launch_mlp_grad(B.X, B.T, nullptr, (int)count, B.W1b, B.P, B.G, B.loss, s);
    if (rc == DTWJ_OK) {
        std::vector<float> P(np);
        float l = 0.f;
        hipError_t e = hipMemcpyAsync(P.data(), B.P, sizeof(float) * np, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(&l, B.loss, sizeof(float), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            rc = hip_fail(e, "trained parameters", __FILE__, __LINE__);
        } else {
            params_to_mlp(P.data(), *model);
            if (final_loss) *final_loss = l;
        }
    }
    dtwj_context_destroy(c);
    return rc;
}

int dtwj_top1_learned(const double *series, int64_t n, int32_t m, int32_t w, const dtwj_mlp *model,
                      int32_t n_items, int32_t n_verify, const dtwj_options *opt, dtwj_result *out,
                      dtwj_proposal_report *report) {
    const auto te = clk::now();
    if (!model || !out) return fail(DTWJ_EINVAL, "null model or result");
    if (int rc = validate(series, n, m, w)) return rc;
    for (int64_t t = 0; t < n; ++t)
        if (!std::isfinite(series[t])) return fail(DTWJ_EINVAL, "series contains NaN or Inf");
    if (n_verify <= 0) n_verify = 1024;
    dtwj_context *c = nullptr;
    if (int rc = with_context(opt ? opt->device : 0, &c)) return rc;
    const double *d = nullptr;
    Proposal Pr;
    dtwj_proposal_report rep;
    std::memset(&rep, 0, sizeof(rep));
    int rc = dtwj_context_upload(c, series, n, &d);
    if (rc == DTWJ_OK) rc = propose(c, d, n, m, w, n_items, n_verify, Pr);
    const int cnt = (int)Pr.cands.size();
    std::vector<int2> learned, by_ed;
    if (rc == DTWJ_OK && cnt > 0) {
        rc = score(c, *model, Pr.d_F, cnt, Pr);
        // rank: log ED distance + predicted log(DTW / ED) (keys computed on the GPU)
        if (rc == DTWJ_OK) rc = launch_prop_keys(Pr.d_F, Pr.d_y, cnt, m, Pr.d_key, Pr.d_key + cnt, c->stream);
        std::vector<float> key(cnt), ed(cnt);
        if (rc == DTWJ_OK) {
            hipStream_t s = c->stream;
            hipError_t e = hipMemcpyAsync(key.data(), Pr.d_key, sizeof(float) * cnt, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(ed.data(), Pr.d_key + cnt, sizeof(float) * cnt, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) rc = hip_fail(e, "proposal scores", __FILE__, __LINE__);
        }
        if (rc == DTWJ_OK) {
            const int take = std::min(n_verify, cnt);
            std::vector<int> order(cnt);
            std::iota(order.begin(), order.end(), 0);
            std::partial_sort(order.begin(), order.begin() + take, order.end(),
                              [&](int a, int b) { return key[a] < key[b]; });
            for (int k = 0; k < take; ++k) learned.push_back(Pr.cands[order[k]]);
            std::iota(order.begin(), order.end(), 0);
            std::partial_sort(order.begin(), order.begin() + take, order.end(),
                              [&](int a, int b) { return ed[a] < ed[b]; });
            for (int k = 0; k < take; ++k) by_ed.push_back(Pr.cands[order[k]]);
            std::vector<double> cl, ce;
            rc = verify_pairs(c, Pr.J, learned, Pr, cl);
            if (rc == DTWJ_OK) rc = verify_pairs(c, Pr.J, by_ed, Pr, ce);
            if (rc == DTWJ_OK) {
                rep.seed_dist = std::sqrt(*std::min_element(cl.begin(), cl.end()));
                rep.seed_dist_ed = std::sqrt(*std::min_element(ce.begin(), ce.end()));
            }
        }
    }
    rep.candidates = cnt;
    rep.verified = (int32_t)learned.size();
    rep.t_propose = secs(te, clk::now());
    if (rc == DTWJ_OK) {
        // the plain join, seeded with the verified proposals (its own ED scan replaced)
        Seeds seeds;
        seeds.pairs = learned;
        seeds.gpu_scan = learned.empty();
        rc = run_join(c, d, n, m, w, opt, seeds, out, te);
    }
    if (rc == DTWJ_OK) {
        if (rep.seed_dist > 0) {
            rep.regret = rep.seed_dist / out->dist - 1.0;
            rep.regret_ed = rep.seed_dist_ed / out->dist - 1.0;
        }
        if (report) *report = rep;
    }
    dtwj_context_destroy(c);
    return rc;
}

// ---------------------------------------------------------------- building blocks
static int building_block_prep(const double *series, int64_t n, int32_t m, int32_t w, dtwj_context **cout,
                               Prep &P, const double **d_a, int32_t paa_seg = 0) {
    if (int rc = with_context(0, cout)) return rc;
    dtwj_context *c = *cout;
    if (int rc = dtwj_context_upload(c, series, n, d_a)) return rc;
    Job J;
    plan_job(J, *d_a, n, m, w, 0, 1, false, {}, 1);
    J.w = w;
    J.paa_seg = paa_seg;
    Arena dry;
    take_job(dry, J);
    if (int rc = ensure_mem(c, dry.off + 256)) return rc;
    Arena ar;
    ar.base = (char *)c->mem;
    ar.cap = c->mem_cap;
    take_job(ar, J);
    P = J.P;
    if (int rc = prep_series(P, c->stream)) return rc;
    DTWJ_HIP(hipStreamSynchronize(c->stream));
    return DTWJ_OK;
}

int dtwj_sliding_stats(const double *series, int64_t n, int32_t m, double *mu, double *sigma) {
    if (!series || !mu || !sigma || m < 2 || n < m) return fail(DTWJ_EINVAL, "bad arguments");
    dtwj_context *c = nullptr;
    Prep P;
    const double *d_a = nullptr;
    int rc = building_block_prep(series, n, m, 0, &c, P, &d_a);
    if (rc == DTWJ_OK) {
        const int64_t nw = n - m + 1;
        hipError_t e1 = hipMemcpy(mu, P.mu, sizeof(double) * nw, hipMemcpyDeviceToHost);
        hipError_t e2 = hipMemcpy(sigma, P.sig, sizeof(double) * nw, hipMemcpyDeviceToHost);
        if (e1 != hipSuccess || e2 != hipSuccess) rc = hip_fail(e1 != hipSuccess ? e1 : e2, "hipMemcpy", __FILE__, __LINE__);
    }
    if (c) dtwj_context_destroy(c);
    return rc;
}

int dtwj_envelope(const double *series, int64_t n, int32_t w, double *lower, double *upper) {
    if (!series || !lower || !upper || w < 0 || n < 2) return fail(DTWJ_EINVAL, "bad arguments");
    // a radius of n - 1 already spans the whole series from every position
    if ((int64_t)w > n - 1) w = (int32_t)(n - 1);
    if (w > kMaxEnvelopeRadius) return fail(DTWJ_EINVAL, "envelope radius min(w, n - 1) must be <= 1536");
    dtwj_context *c = nullptr;
    Prep P;
    const double *d_a = nullptr;
    const int32_t m = (int32_t)std::min<int64_t>(n, 2);
    int rc = building_block_prep(series, n, m, w, &c, P, &d_a);
    if (rc == DTWJ_OK) {
        hipError_t e1 = hipMemcpy(lower, P.L, sizeof(double) * n, hipMemcpyDeviceToHost);
        hipError_t e2 = hipMemcpy(upper, P.U, sizeof(double) * n, hipMemcpyDeviceToHost);
        if (e1 != hipSuccess || e2 != hipSuccess) rc = hip_fail(e1 != hipSuccess ? e1 : e2, "hipMemcpy", __FILE__, __LINE__);
    }
    if (c) dtwj_context_destroy(c);
    return rc;
}

int dtwj_dtw_pairs(const double *series, int64_t n, int32_t m, int32_t w, const int64_t *pi, const int64_t *pj,
                   int64_t npairs, int32_t prec, double *cost) {
    if (int rc = validate(series, n, m, w)) return rc;
    if (!pi || !pj || !cost || npairs < 0) return fail(DTWJ_EINVAL, "bad pair arguments");
    if (prec != 32 && prec != 64) return fail(DTWJ_EINVAL, "prec must be 32 or 64");
    const int64_t nw = n - m + 1;
    for (int64_t p = 0; p < npairs; ++p)
        if (pi[p] < 0 || pj[p] < 0 || pi[p] >= nw || pj[p] >= nw) return fail(DTWJ_EINVAL, "pair index out of range");
    if (npairs == 0) return DTWJ_OK;
    dtwj_context *c = nullptr;
    Prep P;
    const double *d_a = nullptr;
    const int32_t we = std::min(w, m - 1);
    int rc = building_block_prep(series, n, m, we, &c, P, &d_a);
    int64_t *d_pi = nullptr, *d_pj = nullptr;
    double *d_cost = nullptr;
    if (rc == DTWJ_OK) {
        int32_t flags[4];
        hipMemcpy(flags, P.flags, sizeof(flags), hipMemcpyDeviceToHost);
        if (flags[1]) rc = fail(DTWJ_EINVAL, "a window of length m is constant (standard deviation 0)");
    }
    if (rc == DTWJ_OK) {
        if (hipMalloc(&d_pi, sizeof(int64_t) * npairs) != hipSuccess || hipMalloc(&d_pj, sizeof(int64_t) * npairs) != hipSuccess ||
            hipMalloc(&d_cost, sizeof(double) * npairs) != hipSuccess)
            rc = fail(DTWJ_ENOMEM, "device allocation failed");
    }
    if (rc == DTWJ_OK) {
        hipMemcpy(d_pi, pi, sizeof(int64_t) * npairs, hipMemcpyHostToDevice);
        hipMemcpy(d_pj, pj, sizeof(int64_t) * npairs, hipMemcpyHostToDevice);
        JoinArgs A;
        std::memset(&A, 0, sizeof(A));
        A.a = d_a;
        A.mu = P.mu;
        A.win = P.win;
        A.ac = P.ac;
        A.mu_c = P.mu_c;
        A.isg = P.isg;
        A.n = n;
        A.nw = nw;
        A.m = m;
        A.w = we;
        rc = launch_dtw_pairs(A, P.sig, d_pi, d_pj, npairs, prec, d_cost, c->stream);
        if (rc == DTWJ_OK) {
            hipError_t e = hipStreamSynchronize(c->stream);
            if (e == hipSuccess) e = hipMemcpy(cost, d_cost, sizeof(double) * npairs, hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = hip_fail(e, "dtw pairs", __FILE__, __LINE__);
        }
    }
    if (d_pi) (void)hipFree(d_pi);
    if (d_pj) (void)hipFree(d_pj);
    if (d_cost) (void)hipFree(d_cost);
    if (c) dtwj_context_destroy(c);
    return rc;
}

int dtwj_lb_petitjean_pairs(const double *series, int64_t n, int32_t m, int32_t w, const int64_t *pi,
                            const int64_t *pj, int64_t npairs, double *lb) {
    if (int rc = validate(series, n, m, w)) return rc;
    if (m > 4096) return fail(DTWJ_EINVAL, "LB_Petitjean pairs need m <= 4096");
    if (!pi || !pj || !lb || npairs < 0) return fail(DTWJ_EINVAL, "bad pair arguments");
    const int64_t nw = n - m + 1;
    for (int64_t p = 0; p < npairs; ++p)
        if (pi[p] < 0 || pj[p] < 0 || pi[p] >= nw || pj[p] >= nw) return fail(DTWJ_EINVAL, "pair index out of range");
    if (npairs == 0) return DTWJ_OK;
    dtwj_context *c = nullptr;
    Prep P;
    const double *d_a = nullptr;
    int rc = building_block_prep(series, n, m, std::min(w, m - 1), &c, P, &d_a);
    int64_t *d_pi = nullptr, *d_pj = nullptr;
    double *d_lb = nullptr;
    if (rc == DTWJ_OK) {
        if (hipMalloc(&d_pi, sizeof(int64_t) * npairs) != hipSuccess || hipMalloc(&d_pj, sizeof(int64_t) * npairs) != hipSuccess ||
            hipMalloc(&d_lb, sizeof(double) * npairs) != hipSuccess)
            rc = fail(DTWJ_ENOMEM, "device allocation failed");
    }
    if (rc == DTWJ_OK) {
        hipError_t e = hipMemcpy(d_pi, pi, sizeof(int64_t) * npairs, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_pj, pj, sizeof(int64_t) * npairs, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
    }
    if (rc == DTWJ_OK) rc = launch_lb_petitjean(d_a, P.mu, P.sig, m, std::min(w, m - 1), d_pi, d_pj, npairs, d_lb, c->stream);
    if (rc == DTWJ_OK) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(lb, d_lb, sizeof(double) * npairs, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "LB_Petitjean pairs", __FILE__, __LINE__);
    }
    if (d_pi) (void)hipFree(d_pi);
    if (d_pj) (void)hipFree(d_pj);
    if (d_lb) (void)hipFree(d_lb);
    dtwj_context_destroy(c);
    return rc;
}

int dtwj_paa_pairs(const double *series, int64_t n, int32_t m, int32_t w, const int64_t *pi, const int64_t *pj,
                   int64_t npairs, double *lb) {
    if (int rc = validate(series, n, m, w)) return rc;
    if (m < 16) return fail(DTWJ_EINVAL, "the segment-mean bound needs m >= 16");
    if (!pi || !pj || !lb || npairs < 0) return fail(DTWJ_EINVAL, "bad pair arguments");
    const int64_t nw = n - m + 1;
    for (int64_t p = 0; p < npairs; ++p)
        if (pi[p] < 0 || pj[p] < 0 || pi[p] >= nw || pj[p] >= nw) return fail(DTWJ_EINVAL, "pair index out of range");
    if (npairs == 0) return DTWJ_OK;
    dtwj_context *c = nullptr;
    Prep P;
    const double *d_a = nullptr;
    const int32_t nseg = paa_segments(m);
    int rc = building_block_prep(series, n, m, std::min(w, m - 1), &c, P, &d_a, nseg);
    int64_t *d_pi = nullptr, *d_pj = nullptr;
    double *d_lb = nullptr;
    if (rc == DTWJ_OK) {
        if (hipMalloc(&d_pi, sizeof(int64_t) * npairs) != hipSuccess || hipMalloc(&d_pj, sizeof(int64_t) * npairs) != hipSuccess ||
            hipMalloc(&d_lb, sizeof(double) * npairs) != hipSuccess)
            rc = fail(DTWJ_ENOMEM, "device allocation failed");
    }
    if (rc == DTWJ_OK) {
        hipError_t e = hipMemcpy(d_pi, pi, sizeof(int64_t) * npairs, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_pj, pj, sizeof(int64_t) * npairs, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
    }
    if (rc == DTWJ_OK) rc = launch_paa_pairs(P.paa, nseg, m / nseg, d_pi, d_pj, npairs, d_lb, c->stream);
    if (rc == DTWJ_OK) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(lb, d_lb, sizeof(double) * npairs, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "segment-mean pairs", __FILE__, __LINE__);
    }
    if (d_pi) (void)hipFree(d_pi);
    if (d_pj) (void)hipFree(d_pj);
    if (d_lb) (void)hipFree(d_lb);
    if (c) dtwj_context_destroy(c);
    return rc;
}

// The global occupancy mask of a preprocessed series (P: statistics, envelope of radius w) into
// buffers of one device allocation (*mem, freed by the caller).
static int run_global_mask(dtwj_context *c, const Prep &P, int64_t n, int32_t m, int32_t w, MaskBuffers &B,
                           void **mem) {
    B.a = P.a;
    B.U = P.U;
    B.L = P.L;
    B.mu = P.mu;
    B.sig = P.sig;
    B.n = n;
    B.m = m;
    B.w = w;
    B.lmax = mask_table_bins(m);
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t b_table = up(sizeof(uint32_t) * (size_t)B.lmax * m), b_diff = up(sizeof(double) * (size_t)n),
                 b_mask = up((size_t)n), b_small = 256;
    char *base = nullptr;
    if (hipMalloc(&base, b_table + b_diff + b_mask + b_small) != hipSuccess)
        return fail(DTWJ_ENOMEM, "device allocation failed (global mask)");
    *mem = base;
    B.table = reinterpret_cast<uint32_t *>(base);
    B.diff = reinterpret_cast<double *>(base + b_table);
    B.mask = reinterpret_cast<uint8_t *>(base + b_table + b_diff);
    char *small = base + b_table + b_diff + b_mask;
    B.ths = reinterpret_cast<double *>(small);
    B.zmax_bits = reinterpret_cast<unsigned long long *>(small + 8);
    B.status = reinterpret_cast<int32_t *>(small + 16);
    if (int rc = launch_global_mask(B, c->stream)) return rc;
    DTWJ_HIP(hipStreamSynchronize(c->stream));
    int32_t status[2];
    DTWJ_HIP(hipMemcpy(status, B.status, sizeof(status), hipMemcpyDeviceToHost));
    if (status[1]) return fail(DTWJ_EINVAL, "a z-value exceeds the histogram range (window statistics out of range)");
    return DTWJ_OK;
}

static int mask_checks(const double *series, int64_t n, int32_t m, int32_t w) {
    if (int rc = validate(series, n, m, w)) return rc;
    if (w > kMaskMaxRadius) return fail(DTWJ_EINVAL, "the global mask needs w <= 256");
    if (m > kMaskMaxWindow) return fail(DTWJ_EINVAL, "the global mask needs m <= 16384");
    return DTWJ_OK;
}

// Preprocessing of a building block that needs non-constant windows.
static int mask_prep(const double *series, int64_t n, int32_t m, int32_t w, dtwj_context **c, Prep &P) {
    const double *d_a = nullptr;
    int rc = building_block_prep(series, n, m, w, c, P, &d_a);
    if (rc == DTWJ_OK) {
        int32_t flags[4];
        DTWJ_HIP(hipMemcpy(flags, P.flags, sizeof(flags), hipMemcpyDeviceToHost));
        if (flags[1]) rc = fail(DTWJ_EINVAL, "a window of length m is constant (standard deviation 0)");
    }
    return rc;
}

int dtwj_global_mask(const double *series, int64_t n, int32_t m, int32_t w, uint8_t *mask, double *occupancy,
                     double *threshold, int32_t *len_of_table) {
    if (int rc = mask_checks(series, n, m, w)) return rc;
    if (!mask) return fail(DTWJ_EINVAL, "mask is NULL");
    dtwj_context *c = nullptr;
    Prep P;
    void *mem = nullptr;
    MaskBuffers B;
    int rc = mask_prep(series, n, m, w, &c, P);
    if (rc == DTWJ_OK) rc = run_global_mask(c, P, n, m, w, B, &mem);
    if (rc == DTWJ_OK) {
        hipError_t e = hipMemcpy(mask, B.mask, (size_t)n, hipMemcpyDeviceToHost);
        if (e == hipSuccess && occupancy) e = hipMemcpy(occupancy, B.diff, sizeof(double) * n, hipMemcpyDeviceToHost);
        if (e == hipSuccess && threshold) e = hipMemcpy(threshold, B.ths, sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess && len_of_table) e = hipMemcpy(len_of_table, B.status, sizeof(int32_t), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "global mask", __FILE__, __LINE__);
    }
    if (mem) (void)hipFree(mem);
    if (c) dtwj_context_destroy(c);
    return rc;
}

int dtwj_masked_ball_pairs(const double *series, int64_t n, int32_t m, int32_t w, const int64_t *pi,
                           const int64_t *pj, int64_t npairs, double *lb) {
    if (int rc = mask_checks(series, n, m, w)) return rc;
    if (!pi || !pj || !lb || npairs < 0) return fail(DTWJ_EINVAL, "bad pair arguments");
    const int64_t nw = n - m + 1;
    for (int64_t p = 0; p < npairs; ++p)
        if (pi[p] < 0 || pj[p] < 0 || pi[p] >= nw || pj[p] >= nw) return fail(DTWJ_EINVAL, "pair index out of range");
    if (npairs == 0) return DTWJ_OK;
    dtwj_context *c = nullptr;
    Prep P;
    void *mem = nullptr;
    MaskBuffers B;
    int64_t *d_pi = nullptr, *d_pj = nullptr;
    double *d_lb = nullptr;
    int rc = mask_prep(series, n, m, w, &c, P);
    if (rc == DTWJ_OK) rc = run_global_mask(c, P, n, m, w, B, &mem);
    if (rc == DTWJ_OK) {
        if (hipMalloc(&d_pi, sizeof(int64_t) * npairs) != hipSuccess || hipMalloc(&d_pj, sizeof(int64_t) * npairs) != hipSuccess ||
            hipMalloc(&d_lb, sizeof(double) * 2 * npairs) != hipSuccess)
            rc = fail(DTWJ_ENOMEM, "device allocation failed");
    }
    if (rc == DTWJ_OK) {
        hipError_t e = hipMemcpy(d_pi, pi, sizeof(int64_t) * npairs, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_pj, pj, sizeof(int64_t) * npairs, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
    }
    if (rc == DTWJ_OK) rc = launch_masked_ball(B, d_pi, d_pj, npairs, d_lb, c->stream);
    if (rc == DTWJ_OK) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(lb, d_lb, sizeof(double) * 2 * npairs, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "masked ball pairs", __FILE__, __LINE__);
    }
    if (d_pi) (void)hipFree(d_pi);
    if (d_pj) (void)hipFree(d_pj);
    if (d_lb) (void)hipFree(d_lb);
    if (mem) (void)hipFree(mem);
    if (c) dtwj_context_destroy(c);
    return rc;
}

}  // extern "C"
