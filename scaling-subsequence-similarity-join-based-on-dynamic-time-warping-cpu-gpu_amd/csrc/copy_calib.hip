// copy_calib: a plain device-to-device copy of a known size, used to calibrate the HBM
// traffic counters (FETCH_SIZE / WRITE_SIZE) that the join kernel's traffic figure is
// scaled by.  Each launch reads exactly `bytes` and writes exactly `bytes`.
//
// Usage: copy_calib [MiB]   (default 1024)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy_calib(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int main(int argc, char **argv) {
    size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1024;
    if (mib == 0 || mib > 65536) {
        fprintf(stderr, "copy_calib: size must be 1..65536 MiB\n");
        return 2;
    }
    size_t bytes = mib << 20, n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        fprintf(stderr, "copy_calib: allocation failed\n");
        return 1;
    }
    if (hipMemset(a, 1, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
    for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(k_copy_calib, dim3(4096), dim3(256), 0, 0, a, b, n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        fprintf(stderr, "copy_calib: kernel failed\n");
        return 1;
    }
    printf("k_copy_calib: %zu bytes read and %zu bytes written per launch, 3 launches\n", bytes, bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
