// Deterministic synthetic series generator (host code, part of libdtwj).
//
// The workloads the reference is benchmarked on are seeded random walks with planted,
// time-warped motif pairs.  The reference ships no generator (it reads a text file,
// CPU_version1/loadfile.cpp:14-33), so the procedure is fully specified here:
//
//   RNG      : splitmix64(seed) -> 4 words of xoshiro256** state;
//              uniform u in [0,1) = (next() >> 11) * 2^-53;
//              normals by Box-Muller: r = sqrt(-2 ln(1-u1)), z0 = r cos(2 pi u2), z1 = r sin(2 pi u2).
//   walk     : a[0] = s0*N0, a[t] = a[t-1] + s_t*N_t, with s_t = step_std, or, when
//              regime_len > 0, alternating step_std / step_std_alt every regime_len samples.
//   plant p  : source window [src, src+L) with L = round(m*(1+e)), e ~ U(-warp, +warp),
//              resampled to m points by linear interpolation (time warp), shifted so that its
//              first point equals a[dst], written to [dst, dst+m), then N(0, noise_std^2)
//              noise is added to those m points; the part of the walk after dst+m is shifted
//              so that the walk keeps its own increments (no jump at dst+m).
//   offsets  : given explicitly, or drawn uniformly with src+m+2m <= dst, all plants
//              pairwise at least m apart (rejection sampling).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/dtwj.h"

namespace {

struct Xoshiro256ss {
    uint64_t s[4];
    static uint64_t splitmix64(uint64_t &x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Xoshiro256ss(uint64_t seed) {
        uint64_t x = seed;
        for (int k = 0; k < 4; ++k) s[k] = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

struct Normal {
    Xoshiro256ss rng;
    bool have = false;
    double spare = 0.0;
    explicit Normal(uint64_t seed) : rng(seed) {}
    double operator()() {
        if (have) { have = false; return spare; }
        const double u1 = rng.uniform(), u2 = rng.uniform();
        const double r = std::sqrt(-2.0 * std::log(1.0 - u1));
        const double th = 6.283185307179586476925286766559 * u2;
        spare = r * std::sin(th);
        have = true;
        return r * std::cos(th);
    }
};

}  // namespace

extern "C" void dtwj_gen_spec_default(dtwj_gen_spec *spec) {
    std::memset(spec, 0, sizeof(*spec));
    spec->n = 16384;
    spec->seed = 0;
    spec->step_std = 1.0;
    spec->step_std_alt = 0.0;
    spec->regime_len = 0;
    spec->n_plants = 0;
    spec->plant_len = 128;
    spec->warp = 0.05;
    spec->noise_std = 0.1;
    spec->plant_src = nullptr;
    spec->plant_dst = nullptr;
}

extern "C" int dtwj_generate(const dtwj_gen_spec *spec, double *out, int64_t *src_out, int64_t *dst_out) {
    if (!spec || !out || spec->n < 2 || spec->n_plants < 0 || spec->plant_len < 4) return DTWJ_EINVAL;
    const int64_t n = spec->n;
    const int m = spec->plant_len;
    Normal nrm(spec->seed);
    // Random walk.
    double acc = 0.0;
    for (int64_t t = 0; t < n; ++t) {
        double s = spec->step_std;
        if (spec->regime_len > 0 && ((t / spec->regime_len) & 1)) s = spec->step_std_alt;
        acc += s * nrm();
        out[t] = acc;
    }
    // Plant offsets.
    std::vector<int64_t> src(spec->n_plants), dst(spec->n_plants);
    Xoshiro256ss pos_rng(spec->seed ^ 0xA5A5A5A55A5A5A5Aull);
    for (int p = 0; p < spec->n_plants; ++p) {
        const int L = (int)std::llround(m * (1.0 + spec->warp));  // longest source span
        if (spec->plant_src && spec->plant_dst) {
            src[p] = spec->plant_src[p];
            dst[p] = spec->plant_dst[p];
            if (src[p] < 0 || dst[p] < 0 || src[p] + L + 1 > n || dst[p] + m > n) return DTWJ_EINVAL;
            continue;
        }
        int tries = 0;
        for (;;) {
            if (++tries > 100000) return DTWJ_EINVAL;
            const int64_t span = n - (L + 1) - m;
            if (span <= 3 * m) return DTWJ_EINVAL;
            int64_t s0 = (int64_t)(pos_rng.uniform() * (double)span);
            int64_t d0 = (int64_t)(pos_rng.uniform() * (double)span);
            if (s0 > d0) std::swap(s0, d0);
            if (d0 - s0 < 2 * (int64_t)L) continue;
            bool ok = true;
            for (int q = 0; q < p && ok; ++q) {
                const int64_t a[2] = {s0, d0}, b[2] = {src[q], dst[q]};
                for (int x = 0; x < 2; ++x)
                    for (int y = 0; y < 2; ++y)
                        if (std::llabs(a[x] - b[y]) < 2 * (int64_t)L) ok = false;
            }
            if (!ok) continue;
            src[p] = s0;
            dst[p] = d0;
            break;
        }
    }
    // Plant: resample, shift, add noise, re-join the walk after the planted window.
    std::vector<double> win(m);
    for (int p = 0; p < spec->n_plants; ++p) {
        const double e = (2.0 * pos_rng.uniform() - 1.0) * spec->warp;
        int L = (int)std::llround(m * (1.0 + e));
        if (L < 2) L = 2;
        const int64_t s0 = src[p], d0 = dst[p];
        for (int k = 0; k < m; ++k) {
            const double x = (m == 1) ? 0.0 : (double)k * (double)(L - 1) / (double)(m - 1);
            const int64_t x0 = (int64_t)std::floor(x);
            const double f = x - (double)x0;
            const int64_t i0 = s0 + x0;
            const int64_t i1 = (x0 + 1 < L) ? i0 + 1 : i0;
            win[k] = (1.0 - f) * out[i0] + f * out[i1];
        }
        const double shift = out[d0] - win[0];
        const double old_end = out[d0 + m - 1];
        for (int k = 0; k < m; ++k) out[d0 + k] = win[k] + shift + spec->noise_std * nrm();
        const double delta = out[d0 + m - 1] - old_end;
        for (int64_t t = d0 + m; t < n; ++t) out[t] += delta;
        if (src_out) src_out[p] = s0;
        if (dst_out) dst_out[p] = d0;
    }
    return DTWJ_OK;
}
