// fp16_env_check: the packed-width decode the join uses (fp16_env_decode) against the plain
// conversion and addition, bit for bit, for every FP16 pattern in both halves of the word and a
// set of sample values (zero, signs, tiny and large magnitudes).  NaN results compare as NaN.
//
// Usage: fp16_env_check   (prints the mismatch count; exit status 0 when there is none)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "fp16_env.hpp"

namespace {

constexpr int kNx = 16;

// one thread per (pattern, sample value): 65536 * nx threads
__global__ void k_check(const float *xs, int nx, unsigned long long *bad) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= 65536 * nx) return;
    const uint32_t h = (uint32_t)(t & 0xffff);
    const float x = xs[t >> 16];
    const uint32_t wb = h | ((h ^ 0x1234u) << 16);
    float u0, l0, u1, l1;
    dtwj::fp16_env_decode_ref(x, wb, u0, l0);
    dtwj::fp16_env_decode(x, wb, u1, l1);
    const bool same_u = __float_as_uint(u0) == __float_as_uint(u1) || (u0 != u0 && u1 != u1);
    const bool same_l = __float_as_uint(l0) == __float_as_uint(l1) || (l0 != l0 && l1 != l1);
    if (!(same_u && same_l)) atomicAdd(bad, 1ull);
}

}  // namespace

int main() {
    const float hx[kNx] = {0.f,   1.f,     -1.f,   3.5f,   -1234.567f, 1e-3f,  7.0e5f, -2.5e-4f,
                           100.f, 0.1f,    -0.1f,  65504.f, 1e-30f,    -3e7f,  42.f,   1.5e-8f};
    float *dx = nullptr;
    unsigned long long *dbad = nullptr, hbad = 0;
    if (hipMalloc(&dx, sizeof(hx)) != hipSuccess || hipMalloc(&dbad, sizeof(hbad)) != hipSuccess) {
        fprintf(stderr, "fp16_env_check: allocation failed\n");
        return 2;
    }
    if (hipMemcpy(dx, hx, sizeof(hx), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dbad, &hbad, sizeof(hbad), hipMemcpyHostToDevice) != hipSuccess)
        return 2;
    const int n = 65536 * kNx;
    hipLaunchKernelGGL(k_check, dim3((n + 255) / 256), dim3(256), 0, 0, dx, kNx, dbad);
    if (hipGetLastError() != hipSuccess || hipMemcpy(&hbad, dbad, sizeof(hbad), hipMemcpyDeviceToHost) != hipSuccess) {
        fprintf(stderr, "fp16_env_check: launch failed\n");
        return 2;
    }
    printf("fp16_env_check: %llu mismatches of %d decodes\n", hbad, 2 * n);
    hipFree(dx);
    hipFree(dbad);
    return hbad == 0 ? 0 : 1;
}
