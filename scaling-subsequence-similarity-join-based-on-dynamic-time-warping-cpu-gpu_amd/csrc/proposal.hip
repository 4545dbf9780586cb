// Learned proposal (no reference counterpart; the nearest analogue is the heuristic alpha
// search of CPU_version1/new_dtw_motifGUI.cpp:411-468): candidate features and a 40 -> 64 -> 1
// MLP whose forward and backward passes run on BF16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32
// accumulation.  The network only orders candidate pairs; the pairs it ranks first are
// verified by exact DTW and seed the plain join's best-so-far, so the join stays exact.
//
// MFMA 16x16x16 operand layout (lane l, element e = 0..3, g = l / 16):
//   A[i][k] at lane i + 16 (k / 4), element k % 4;  B[k][j] at lane j + 16 (k / 4), element k % 4;
//   D[i][j] at lane j + 16 (i / 4), element i % 4.
// Forward: M = 16 samples, K = 48 padded features (3 steps), N = 64 hidden (4 tiles).
// Weight gradient dW1 = X^T dH: M = 48 features (3 tiles), K = 16 samples, N = 64 hidden; the
// hidden gradient dH comes out of the forward in the D layout, which is exactly the B layout.
#include "common.hpp"
#include "../../include/dtwj.h"
#include "internal.hpp"

namespace dtwj {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NF = DTWJ_MLP_IN;    // raw features
constexpr int NFP = 48;            // padded to three MFMA k-steps
constexpr int NH = DTWJ_MLP_HID;   // hidden units
constexpr int P_W1 = 0;            // parameter layout: W1t[h][f] (h-major, NFP columns)
constexpr int P_B1 = NH * NFP;
constexpr int P_W2 = P_B1 + NH;
constexpr int P_B2 = P_W2 + NH;
constexpr int P_N = P_B2 + 1;

__host__ __device__ inline uint16_t to_bf16(float f) {
    // round to nearest even (inputs are finite)
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float sqf(float x) { return x * x; }

// ---- candidate features, one thread per candidate pair (i, j):
//  [0,16)  PAA-16 of z_i          [16,32) PAA-16 of z_j
//  32 log1p(LB_Keogh(z_j, env_i)/m)   33 log1p(LB_Keogh(z_i, env_j)/m)
//  34 log1p(ball_i(z_j)/m)            35 log1p(ball_j(z_i)/m)
//  36 log1p(ED^2/m)                   37 log1p(LB_Kim/m)
//  38 (j - i) / n                     39 mean |PAA_i - PAA_j|
__global__ void k_prop_features(JoinArgs A, const int2 *__restrict__ cand, int count, float *__restrict__ F) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= count) return;
    const int i = cand[t].x, j = cand[t].y;
    const int m = A.m;
    const float mci = A.mu_c[i], isi = A.win[i].isig, Ri = A.win[i].R;
    const float mcj = A.mu_c[j], isj = A.win[j].isig, Rj = A.win[j].R;
    float *f = F + (size_t)t * NF;
    float lbA = 0.f, lbB = 0.f, bA = 0.f, bB = 0.f, ed = 0.f, dpaa = 0.f;
    for (int sg = 0; sg < 16; ++sg) {
        const int lo = (int)(((int64_t)sg * m) / 16), hi = (int)(((int64_t)(sg + 1) * m) / 16);
        float pi = 0.f, pj = 0.f;
        for (int k = lo; k < hi; ++k) {
            const float zi = (A.ac[i + k] - mci) * isi, zj = (A.ac[j + k] - mcj) * isj;
            const float Ui = (A.uc[i + k] - mci) * isi, Li = (A.lc[i + k] - mci) * isi;
            const float Uj = (A.uc[j + k] - mcj) * isj, Lj = (A.lc[j + k] - mcj) * isj;
            pi += zi;
            pj += zj;
            lbA += sqf(fmaxf(fmaxf(zj - Ui, Li - zj), 0.f));
            lbB += sqf(fmaxf(fmaxf(zi - Uj, Lj - zi), 0.f));
            bA += sqf(zj - 0.5f * (Ui + Li));
            bB += sqf(zi - 0.5f * (Uj + Lj));
            ed += sqf(zi - zj);
        }
        const float inv = 1.f / (float)(hi - lo);
        f[sg] = pi * inv;
        f[16 + sg] = pj * inv;
        dpaa += fabsf(pi - pj) * inv;
    }
    float zi3[3], zj3[3], zie[3], zje[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        zi3[q] = (A.ac[i + q] - mci) * isi;
        zj3[q] = (A.ac[j + q] - mcj) * isj;
        zie[q] = (A.ac[i + m - 1 - q] - mci) * isi;
        zje[q] = (A.ac[j + m - 1 - q] - mcj) * isj;
    }
    const float im = 1.f / (float)m;
    const float ballA = sqf(fmaxf(sqrtf(bA) - Ri, 0.f)), ballB = sqf(fmaxf(sqrtf(bB) - Rj, 0.f));
    // LB_Kim over the 3x3 corners (every warping path leaves each corner block through one of
    // five cells)
    auto corner = [](const float *x, const float *y) {
        const float c00 = sqf(y[0] - x[0]), c10 = sqf(y[1] - x[0]), c01 = sqf(y[0] - x[1]), c11 = sqf(y[1] - x[1]);
        const float c20 = sqf(y[2] - x[0]), c21 = sqf(y[2] - x[1]), c22 = sqf(y[2] - x[2]), c12 = sqf(y[1] - x[2]),
                    c02 = sqf(y[0] - x[2]);
        float b = c10 + c20;
        b = fminf(b, fminf(c10, c11) + c21);
        b = fminf(b, c11 + c22);
        b = fminf(b, fminf(c01, c11) + c12);
        b = fminf(b, c01 + c02);
        return c00 + b;
    };
    const float kim = corner(zi3, zj3) + corner(zie, zje);
    f[32] = log1pf(lbA * im);
    f[33] = log1pf(lbB * im);
    f[34] = log1pf(ballA * im);
    f[35] = log1pf(ballB * im);
    f[36] = log1pf(ed * im);
    f[37] = log1pf(kim * im);
    f[38] = (float)(j - i) / (float)A.n;
    f[39] = dpaa * (1.f / 16.f);
}

// ---- (x - mean) * scale, BF16, padded to NFP columns and to a multiple of 16 rows
__global__ void k_prop_pack(const float *__restrict__ F, int count, int rows, const float *__restrict__ mean,
                            const float *__restrict__ scale, uint16_t *__restrict__ X) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= rows * NFP) return;
    const int r = t / NFP, c = t - r * NFP;
    float v = 0.f;
    if (r < count && c < NF) v = (F[(size_t)r * NF + c] - mean[c]) * scale[c];
    X[t] = to_bf16(v);
}

// Loads 4 consecutive BF16 values (8-byte aligned) as an MFMA operand.
__device__ inline bf16x4 load4(const uint16_t *p) {
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    bf16x4 r;
    r[0] = (short)(u.x & 0xffffu);
    r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xffffu);
    r[3] = (short)(u.y >> 16);
    return r;
}

// Forward of one 16-sample tile: arow is the row of X holding sample l % 16 of the tile (lane
// l's A operand).  Returns the hidden activations (D layout) and the outputs of samples 4g + e
// (on every lane of group g).
__device__ inline void mlp_tile_forward(const uint16_t *__restrict__ X, int arow, const uint16_t *__restrict__ W1b,
                                        const float *__restrict__ P, f32x4 H[4], float y[4]) {
    const int l = (int)threadIdx.x & 63;
    const int g = l >> 4, c = l & 15;
    const uint16_t *xr = X + (size_t)arow * NFP;
    bf16x4 a[3];
#pragma unroll
    for (int ks = 0; ks < 3; ++ks) a[ks] = load4(xr + 16 * ks + 4 * g);
#pragma unroll
    for (int ht = 0; ht < 4; ++ht) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        const uint16_t *wr = W1b + (size_t)(16 * ht + c) * NFP;
#pragma unroll
        for (int ks = 0; ks < 3; ++ks) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ks], load4(wr + 16 * ks + 4 * g), acc, 0, 0, 0);
        const float b1 = P[P_B1 + 16 * ht + c];
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[e] = fmaxf(acc[e] + b1, 0.f);
        H[ht] = acc;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        float s = 0.f;
#pragma unroll
        for (int ht = 0; ht < 4; ++ht) s = fmaf(H[ht][e], P[P_W2 + 16 * ht + c], s);
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 16);
        y[e] = s + P[P_B2];
    }
}

__global__ __launch_bounds__(64) void k_mlp_forward(const uint16_t *__restrict__ X, int count,
                                                    const uint16_t *__restrict__ W1b, const float *__restrict__ P,
                                                    float *__restrict__ out) {
    const int l = (int)threadIdx.x;
    const int ntile = (count + 15) / 16;
    for (int tile = (int)blockIdx.x; tile < ntile; tile += (int)gridDim.x) {
        f32x4 H[4];
        float y[4];
        mlp_tile_forward(X, tile * 16 + (l & 15), W1b, P, H, y);  // X is padded to whole tiles
        if ((l & 15) == 0) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int s = tile * 16 + 4 * (l >> 4) + e;
                if (s < count) out[s] = y[e];
            }
        }
    }
}

// Gradient of the mean squared error over the batch samples idx[0..bsz) (idx NULL: 0..bsz),
// accumulated into G (parameter layout) and *loss; waves stride over 16-sample tiles.
__global__ __launch_bounds__(64) void k_mlp_grad(const uint16_t *__restrict__ X, const float *__restrict__ target,
                                                 const int32_t *__restrict__ idx, int bsz,
                                                 const uint16_t *__restrict__ W1b, const float *__restrict__ P,
                                                 float *__restrict__ G, float *__restrict__ loss) {
    const int l = (int)threadIdx.x;
    const int g = l >> 4, c = l & 15;
    const float inv_b = 1.f / (float)bsz;
    f32x4 gW[3][4];
#pragma unroll
    for (int ft = 0; ft < 3; ++ft)
#pragma unroll
        for (int ht = 0; ht < 4; ++ht) gW[ft][ht] = f32x4{0.f, 0.f, 0.f, 0.f};
    float gb1[4] = {0.f, 0.f, 0.f, 0.f}, gw2[4] = {0.f, 0.f, 0.f, 0.f};
    float gb2 = 0.f, lsum = 0.f;
    const int ntile = (bsz + 15) / 16;
    auto row_of = [&](int b) { return b < bsz ? (idx ? idx[b] : b) : (idx ? idx[0] : 0); };
    for (int tile = (int)blockIdx.x; tile < ntile; tile += (int)gridDim.x) {
        // samples 4g + e of this lane's group (rows of X), and their validity
        int srow[4];
        bool sval[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int b = tile * 16 + 4 * g + e;
            sval[e] = b < bsz;
            srow[e] = row_of(b);
        }
        f32x4 H[4];
        float y[4];
        mlp_tile_forward(X, row_of(tile * 16 + c), W1b, P, H, y);
        float dy[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float d = sval[e] ? y[e] - target[srow[e]] : 0.f;
            dy[e] = 2.f * d * inv_b;
            if (c == 0) lsum += d * d * inv_b;
            if (c == 0) gb2 += dy[e];
        }
        // X^T operand of the weight gradient: A[f][s] = X[s][f], s = 4g + e
        bf16x4 xt[3];
#pragma unroll
        for (int ft = 0; ft < 3; ++ft)
#pragma unroll
            for (int e = 0; e < 4; ++e) xt[ft][e] = (short)X[(size_t)srow[e] * NFP + 16 * ft + c];
#pragma unroll
        for (int ht = 0; ht < 4; ++ht) {
            const float w2 = P[P_W2 + 16 * ht + c];
            bf16x4 dh;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float d = H[ht][e] > 0.f ? dy[e] * w2 : 0.f;
                dh[e] = (short)to_bf16(d);
                gb1[ht] += __builtin_bit_cast(float, (uint32_t)(uint16_t)dh[e] << 16);
                gw2[ht] += dy[e] * H[ht][e];
            }
#pragma unroll
            for (int ft = 0; ft < 3; ++ft) gW[ft][ht] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xt[ft], dh, gW[ft][ht], 0, 0, 0);
        }
    }
    // gW[ft][ht] element e: dW1[f = 16 ft + 4 g + e][h = 16 ht + c]
#pragma unroll
    for (int ft = 0; ft < 3; ++ft)
#pragma unroll
        for (int ht = 0; ht < 4; ++ht)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int f = 16 * ft + 4 * g + e, h = 16 * ht + c;
                if (f < NF) atomicAdd(&G[P_W1 + h * NFP + f], gW[ft][ht][e]);
            }
#pragma unroll
    for (int ht = 0; ht < 4; ++ht) {
        float a = gb1[ht], b = gw2[ht];
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 32, 64);
        b += __shfl_xor(b, 16, 64);
        b += __shfl_xor(b, 32, 64);
        if (g == 0) {
            atomicAdd(&G[P_B1 + 16 * ht + c], a);
            atomicAdd(&G[P_W2 + 16 * ht + c], b);
        }
    }
    // gb2 and the loss live on lanes with c == 0 (one per group)
    gb2 += __shfl_xor(gb2, 16, 64);
    gb2 += __shfl_xor(gb2, 32, 64);
    lsum += __shfl_xor(lsum, 16, 64);
    lsum += __shfl_xor(lsum, 32, 64);
    if (l == 0) {
        atomicAdd(&G[P_B2], gb2);
        atomicAdd(loss, lsum);
    }
}

// Adam on the FP32 master parameters; refreshes the BF16 copy of W1.
__global__ void k_mlp_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ M1,
                           float *__restrict__ M2, float lr, float bc1, float bc2, uint16_t *__restrict__ W1b) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= P_N) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gr = G[t];
    const float m1 = b1 * M1[t] + (1.f - b1) * gr;
    const float m2 = b2 * M2[t] + (1.f - b2) * gr * gr;
    M1[t] = m1;
    M2[t] = m2;
    const bool pad = t < P_B1 && (t % NFP) >= NF;
    const float p = pad ? 0.f : P[t] - lr * (m1 / bc1) / (sqrtf(m2 / bc2) + eps);
    P[t] = p;
    if (t < P_B1) W1b[t] = to_bf16(p);
}

// Ranking keys of the candidates: log ED distance + network output, and log ED distance
// alone (ED^2 = m expm1(feature 36)).
__global__ void k_prop_keys(const float *__restrict__ F, const float *__restrict__ y, int count, int m,
                            float *__restrict__ key, float *__restrict__ key_ed) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= count) return;
    const float led = 0.5f * logf((float)m * expm1f(F[(size_t)t * NF + 36]) + 1e-30f);
    key[t] = led + y[t];
    key_ed[t] = led;
}

__global__ void k_fill_f32(float *p, int n, float v) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t < n) p[t] = v;
}

__global__ void k_params_to_bf16(const float *__restrict__ P, uint16_t *__restrict__ W1b) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t < P_B1) W1b[t] = to_bf16(P[t]);
}

}  // namespace

int mlp_param_count() { return P_N; }
int mlp_feature_pad() { return NFP; }

// Host layout conversion: dtwj_mlp <-> device parameter vector.
void mlp_to_params(const dtwj_mlp &M, float *P) {
    for (int i = 0; i < P_N; ++i) P[i] = 0.f;
    for (int h = 0; h < NH; ++h) {
        for (int f = 0; f < NF; ++f) P[P_W1 + h * NFP + f] = M.w1[h][f];
        P[P_B1 + h] = M.b1[h];
        P[P_W2 + h] = M.w2[h];
    }
    P[P_B2] = M.b2;
}

void params_to_mlp(const float *P, dtwj_mlp &M) {
    for (int h = 0; h < NH; ++h) {
        for (int f = 0; f < NF; ++f) M.w1[h][f] = P[P_W1 + h * NFP + f];
        M.b1[h] = P[P_B1 + h];
        M.w2[h] = P[P_W2 + h];
    }
    M.b2 = P[P_B2];
}

// Candidate pairs as 64-bit keys (i << 32 | j, both non-negative and below 2^31), entries
// without a pair, and the padding up to the sorted length, as the largest key, so that an
// ascending sort orders pairs by (i, j) and puts the empty entries last.
__global__ __launch_bounds__(256) void k_pair_keys(const int2 *__restrict__ cand, int64_t count, int64_t len,
                                                   uint64_t *key) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= len) return;
    uint64_t k = ~0ull;
    if (t < count) {
        const int2 p = cand[t];
        if (p.x >= 0) k = ((uint64_t)(uint32_t)p.x << 32) | (uint32_t)p.y;
    }
    key[t] = k;
}

// Bitonic sorting network over len = 2^p keys.  Stage k (the length of the runs being merged,
// 2 .. len) takes steps j = k / 2 .. 1; step j compares key i with key i + j for every i whose
// bit j is clear, and orders the two ascending where bit k of i is clear, descending where it is
// set.  Thread t of a step owns the pair whose lower index is i = 2 t - (t mod j), so every key
// is touched once per step.  Steps with j < SB / 2 stay inside one block of SB keys and run in
// LDS, all of them in one launch; the wider steps of a stage run one launch each over HBM.
constexpr int SB = 2048;  // keys per block of the LDS steps (1024 threads, one pair each)

__device__ __forceinline__ void cmp_swap(uint64_t &a, uint64_t &b, bool up) {
    if ((a > b) == up) {
        const uint64_t t = a;
        a = b;
        b = t;
    }
}

// Stages k_lo .. k_hi (powers of two), each from step min(k, SB) / 2 down to 1, over the block's
// SB keys; the direction of a pair follows the global index of its lower key.
__global__ __launch_bounds__(1024) void k_bitonic_lds(uint64_t *key, int64_t k_lo, int64_t k_hi) {
    __shared__ uint64_t v[SB];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * SB;
    v[t] = key[base + t];
    v[t + SB / 2] = key[base + t + SB / 2];
    __syncthreads();
    for (int64_t k = k_lo; k <= k_hi; k <<= 1) {
        for (int j = (int)(k < SB ? k : SB) / 2; j > 0; j >>= 1) {
            const int i = 2 * t - (t & (j - 1));
            cmp_swap(v[i], v[i + j], ((base + i) & k) == 0);
            __syncthreads();
        }
    }
    key[base + t] = v[t];
    key[base + t + SB / 2] = v[t + SB / 2];
}

// One step j >= SB / 2 of stage k over all len keys (len / 2 threads).
__global__ __launch_bounds__(256) void k_bitonic_step(uint64_t *key, int64_t half, int64_t k, int64_t j) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= half) return;
    const int64_t i = 2 * t - (t & (j - 1));
    uint64_t a = key[i], b = key[i + j];
    const bool up = (i & k) == 0;
    if ((a > b) == up) {
        key[i] = b;
        key[i + j] = a;
    }
}

int64_t pair_sort_len(int64_t count) {
    int64_t len = SB;
    while (len < count) len <<= 1;
    return len;
}

int launch_pair_sort(const int2 *cand, int64_t count, uint64_t *key, int64_t len, hipStream_t s) {
    if (count <= 0) return 0;
    if (len < count || len < SB || (len & (len - 1)) != 0) {
        set_error("pair sort: the key buffer length must be a power of two, at least 2048 and the count");
        return DTWJ_EINVAL;
    }
    hipLaunchKernelGGL(k_pair_keys, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, cand, count, len, key);
    DTWJ_HIP(hipGetLastError());
    const unsigned blocks = (unsigned)(len / SB);
    // stages up to SB entirely in LDS
    hipLaunchKernelGGL(k_bitonic_lds, dim3(blocks), dim3(SB / 2), 0, s, key, (int64_t)2, (int64_t)SB);
    DTWJ_HIP(hipGetLastError());
    for (int64_t k = 2 * (int64_t)SB; k <= len; k <<= 1) {
        for (int64_t j = k / 2; j >= SB; j >>= 1) {
            hipLaunchKernelGGL(k_bitonic_step, dim3((unsigned)((len / 2 + 255) / 256)), dim3(256), 0, s, key, len / 2,
                               k, j);
            DTWJ_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_bitonic_lds, dim3(blocks), dim3(SB / 2), 0, s, key, k, k);
        DTWJ_HIP(hipGetLastError());
    }
    return 0;
}

int launch_prop_features(const JoinArgs &A, const int2 *cand, int count, float *F, hipStream_t s) {
    if (count <= 0) return 0;
    hipLaunchKernelGGL(k_prop_features, dim3((unsigned)((count + 127) / 128)), dim3(128), 0, s, A, cand, count, F);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_prop_pack(const float *F, int count, int rows, const float *mean, const float *scale, uint16_t *X,
                     hipStream_t s) {
    const int total = rows * NFP;
    hipLaunchKernelGGL(k_prop_pack, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, F, count, rows, mean, scale, X);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_mlp_forward(const uint16_t *X, int count, const uint16_t *W1b, const float *P, float *out, hipStream_t s) {
    const int ntile = (count + 15) / 16;
    const int grid = ntile < 2048 ? (ntile > 0 ? ntile : 1) : 2048;
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(64), 0, s, X, count, W1b, P, out);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_mlp_grad(const uint16_t *X, const float *target, const int32_t *idx, int bsz, const uint16_t *W1b,
                    const float *P, float *G, float *loss, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_f32, dim3((P_N + 1 + 255) / 256), dim3(256), 0, s, G, P_N, 0.f);
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, s, loss, 1, 0.f);
    const int ntile = (bsz + 15) / 16;
    const int grid = ntile < 512 ? (ntile > 0 ? ntile : 1) : 512;
    hipLaunchKernelGGL(k_mlp_grad, dim3(grid), dim3(64), 0, s, X, target, idx, bsz, W1b, P, G, loss);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_mlp_adam(float *P, const float *G, float *M1, float *M2, float lr, int step, uint16_t *W1b, hipStream_t s) {
    const float bc1 = 1.f - powf(0.9f, (float)step), bc2 = 1.f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((P_N + 255) / 256), dim3(256), 0, s, P, G, M1, M2, lr, bc1, bc2, W1b);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_prop_keys(const float *F, const float *y, int count, int m, float *key, float *key_ed, hipStream_t s) {
    if (count <= 0) return 0;
    hipLaunchKernelGGL(k_prop_keys, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, F, y, count, m, key, key_ed);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

int launch_params_to_bf16(const float *P, uint16_t *W1b, hipStream_t s) {
    hipLaunchKernelGGL(k_params_to_bf16, dim3((P_B1 + 255) / 256), dim3(256), 0, s, P, W1b);
    DTWJ_HIP(hipGetLastError());
    return 0;
}

}  // namespace dtwj
