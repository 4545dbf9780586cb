// Shared definitions for the gfx950 join: error plumbing, double-double arithmetic,
// wave-level helpers and the argument blocks passed from the host orchestration
// (dtwj_api.cpp) to the kernels (prep.hip, join.hip, dtw.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#define DTWJ_WAVE 64

// ---------------------------------------------------------------- host error plumbing
namespace dtwj {
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);
}  // namespace dtwj

#define DTWJ_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) return dtwj::hip_fail(e_, #call, __FILE__, __LINE__);     \
    } while (0)

// ---------------------------------------------------------------- double-double (FP64x2)
// An unevaluated sum hi + lo with |lo| <= ulp(hi)/2.  Used for the compensated prefix sums
// of the sliding statistics (Knuth TwoSum / FMA TwoProd).
struct dd_t {
    double hi, lo;
};

// Error-free transformations must not be contracted: with -ffp-contract=fast the compiler
// would fuse a caller's product into TwoSum's addition and silently drop the error term.
#define DTWJ_NO_CONTRACT _Pragma("clang fp contract(off)")

__host__ __device__ inline dd_t dd_two_sum(double a, double b) {
    DTWJ_NO_CONTRACT
    double s = a + b;
    double bb = s - a;
    double err = (a - (s - bb)) + (b - bb);
    return {s, err};
}
__host__ __device__ inline dd_t dd_quick_two_sum(double a, double b) {
    DTWJ_NO_CONTRACT
    double s = a + b;
    return {s, b - (s - a)};
}
__host__ __device__ inline dd_t dd_add(dd_t a, dd_t b) {
    DTWJ_NO_CONTRACT
    dd_t s = dd_two_sum(a.hi, b.hi);
    dd_t t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return dd_quick_two_sum(s.hi, s.lo);
}
__host__ __device__ inline dd_t dd_neg(dd_t a) { return {-a.hi, -a.lo}; }
__host__ __device__ inline dd_t dd_sub(dd_t a, dd_t b) { return dd_add(a, dd_neg(b)); }
__host__ __device__ inline dd_t dd_mul_d(dd_t a, double b) {
    DTWJ_NO_CONTRACT
    double p = a.hi * b;
    double e = fma(a.hi, b, -p);
    e = fma(a.lo, b, e);
    return dd_quick_two_sum(p, e);
}
__host__ __device__ inline dd_t dd_mul(dd_t a, dd_t b) {
    DTWJ_NO_CONTRACT
    double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e = fma(a.hi, b.lo, e);
    e = fma(a.lo, b.hi, e);
    return dd_quick_two_sum(p, e);
}
__host__ __device__ inline dd_t dd_div_d(dd_t a, double b) {
    DTWJ_NO_CONTRACT
    double q1 = a.hi / b;
    dd_t r = dd_sub(a, dd_mul_d(dd_t{q1, 0.0}, b));
    double q2 = r.hi / b;
    return dd_quick_two_sum(q1, q2);
}

// ---------------------------------------------------------------- packed best-so-far
// 64-bit key = (FP32 bits of the squared DTW cost) << 32 | i.  Non-negative floats order
// like their bit patterns, so an unsigned 64-bit atomicMin keeps the smallest cost and,
// among equal costs, the smallest i.
__host__ __device__ inline unsigned long long pack_bsf(float cost, uint32_t i) {
    return ((unsigned long long)__builtin_bit_cast(uint32_t, cost) << 32) | (unsigned long long)i;
}
__host__ __device__ inline float bsf_cost(unsigned long long key) {
    return __builtin_bit_cast(float, (uint32_t)(key >> 32));
}

// ---------------------------------------------------------------- device helpers
#ifdef __HIP_DEVICE_COMPILE__
__device__ inline int lane_id() { return (int)__lane_id(); }
#endif

// Argument blocks -------------------------------------------------------------------------
// Entries of each wave's ball-survivor queue (4 bytes each) in global memory.
constexpr int kJoinQueueCap = 4096;

// Per-window table, 8 floats (32 bytes, two 16-byte vectors).  The first vector holds what
// the one-sided sweep reads for a column (dfa, dga, 1/sigma).
struct WinRow {
    float dfa, dga;  // MPX update terms of the series
    float isig;      // 1 / sigma
    float dfv, dgv;  // MPX update terms of the envelope mid-line v = (U+L)/2
    float mN;        // m + ||(v_i - mu_i)/sigma_i||^2
    float R;         // ball radius ||(U_i - L_i)/2|| / sigma_i
    float sig;       // sigma
};

// Segment-mean (PAA) bound of LB_Keogh: each window is cut into S segments of floor(m / S)
// samples (the rest of the window is left out), S = paa_segments(m).  Row p of the table (3 S
// floats) holds, in the z-units of window p, the segment means of the upper envelope, of the
// lower envelope and of the window itself: [U_0 .. U_S-1, L_0 .. L_S-1, Z_0 .. Z_S-1], FP32, U
// rounded up and L rounded down after widening both by the FP32 rounding bound of any Z.
// 32 segments for windows of 512 and more (measured: m = 512, w = 51 -2.6 %, m = 1024, w = 102
// -4.1 % against 16), 16 below (m = 256, w = 13: +1.7 % with 32).
constexpr int kPaaSegMax = 32;
__host__ __device__ constexpr int paa_segments(int m) { return m >= 512 ? 32 : 16; }

// One series of a join: everything the join kernels read.  Lives in device memory (an
// array of them for a batch); device pointers unless noted.
struct JoinArgs {
    // series and windows
    const double *a;       // n, FP64
    const double *v;       // n, FP64 envelope mid-line
    const double *mu;      // nw, FP64 window means of a
    const double *muv;     // nw, FP64 window means of v
    const WinRow *win;     // nw
    const float *ac;       // n, a - c0 (FP32)
    const float *uc;       // n, U - c0
    const float *lc;       // n, L - c0
    const uint32_t *wc;    // n, envelope widths (U - a, a - L) as FP16 rounded up: lo, hi half
    const float *mu_c;     // nw, mu - c0
    const float *isg;      // nw, 1 / sigma (the WinRow value, as a flat array)
    const float *blkmax;   // 3 x nblk: max sigma, max sv, max 1/sigma over 256-window blocks
    const float *gstat;    // preprocessing statistics (value range, see Prep::gstat)
    const int32_t *flags;  // preprocessing flags (non-finite input, constant window)
    int32_t nblk;
    // problem
    int64_t n, nw;
    int32_t m, w;
    // work decomposition
    const int64_t *item_start;  // G+1 cumulative item counts per diagonal group
    int32_t groups;
    int32_t rows_per_item;
    int64_t n_items;
    // shared best-so-far and results
    unsigned long long *bsf;      // packed
    unsigned long long init_key;  // best-so-far word the set-up kernel starts from (a valid bound)
    int4 *cands;                  // (i, j, cost bits, 0)
    unsigned long long *n_cands;  // pushes so far (64-bit: counts past cand_cap without wrapping)
    int64_t cand_cap;
    unsigned long long *counters; // per-stage counters (see dtwj_api.cpp)
    // error model, filled on the device from gstat by the set-up kernel
    float amax;        // max |ac|, |uc|, |lc|
    float eps_z;       // bound on |z32 - z| for z computed from ac/mu_c/isig
    float rel_dtw;     // relative inflation of the best-so-far cost used as pruning threshold
    int32_t valid;     // 0: the preprocessing flagged the series; the join skips it
    int32_t sweep_mode;  // 0 adaptive, 1 two-sided, 2 one-sided ball test in the sweep
    const float *paa;    // nw x 3 paa_seg segment-mean table, or null: no segment-mean stage
    int32_t paa_seg;     // segments per window (16 or 32; 0 without the stage)
    int32_t paa_len;     // samples per segment, floor(m / paa_seg)
};

// seed scan: per-item best ED pair
struct SeedArgs {
    const double *a;
    const double *mu;
    const WinRow *win;
    int64_t n, nw;
    int32_t m;
    const int64_t *item_start;
    int32_t groups;
    int64_t n_items;
    int32_t rows_per_item;  // rows per join work item (item -> row segment)
    int32_t rows;           // rows scanned per seed item
    int32_t n_seed;         // number of seed items (spread over the item space)
};

// One series of a batched seed scan: its scan arguments and seed-pair buffer; the blocks of all
// series are numbered consecutively (base[t] .. base[t+1] - 1 belong to series t).
struct SeedJob {
    SeedArgs S;
    int2 *pairs;
};

// One series of a batched seed-pair DTW launch: its join descriptor index and seed pairs.
struct SeedDtwJob {
    const int2 *pairs;
    int32_t series;
    int32_t npairs;
};
