// Banded DTW over an explicit pair list, one wavefront per pair.
//  * prec 32: FP32 cells on z = (ac - mu_c) * isig (the join's own arithmetic);
//  * prec 64: FP64 cells on z = (a - mu) / sigma (the re-verification of the join's winners;
//    same formula as the reference's z-normalised subsequences,
//    CPU_version1/new_dtw_motifGUI.cpp:41-44).
#include "dtw_device.hpp"
#include "../../include/dtwj.h"
#include "internal.hpp"

namespace dtwj {

namespace {

template <typename T, int QP>
__global__ __launch_bounds__(64) void k_dtw_pairs(JoinArgs A, const double *__restrict__ sig64,
                                                  const int64_t *__restrict__ pi, const int64_t *__restrict__ pj,
                                                  int64_t npairs, double *cost) {
    extern __shared__ unsigned char smem_raw[];
    T *zi = reinterpret_cast<T *>(smem_raw);
    T *zj = zi + A.m;
    const int64_t p = blockIdx.x;
    if (p >= npairs) return;
    const int64_t i = pi[p], j = pj[p];
    const int lane = threadIdx.x;
    if (sizeof(T) == 8) {
        const double mi = A.mu[i], si = sig64[i], mj = A.mu[j], sj = sig64[j];
        for (int k = lane; k < A.m; k += 64) {
            zi[k] = (T)((A.a[i + k] - mi) / si);
            zj[k] = (T)((A.a[j + k] - mj) / sj);
        }
    } else {
        const float mi = A.mu_c[i], ii = A.win[i].isig, mj = A.mu_c[j], ij = A.win[j].isig;
        for (int k = lane; k < A.m; k += 64) {
            zi[k] = (T)((A.ac[i + k] - mi) * ii);
            zj[k] = (T)((A.ac[j + k] - mj) * ij);
        }
    }
    __syncthreads();
    const PlainZ<T> z{zi, zj};
    int t_stop = 0;
    const T r = dtw_band<T, QP>(A.m, A.w, 64, 1, dev_inf<T>(), z, t_stop);
    if (lane == 0) cost[p] = (double)r;
}

template <typename T>
int launch_pairs_t(const JoinArgs &A, const double *sig64, const int64_t *pi, const int64_t *pj, int64_t npairs,
                   double *cost, hipStream_t s) {
    const int QP = dtw_qp_for(A.w);
    const size_t lds = 2 * (size_t)A.m * sizeof(T);
    const dim3 grid((unsigned)npairs), block(64);
    switch (QP) {
        case 1: hipLaunchKernelGGL((k_dtw_pairs<T, 1>), grid, block, lds, s, A, sig64, pi, pj, npairs, cost); break;
        case 2: hipLaunchKernelGGL((k_dtw_pairs<T, 2>), grid, block, lds, s, A, sig64, pi, pj, npairs, cost); break;
        case 4: hipLaunchKernelGGL((k_dtw_pairs<T, 4>), grid, block, lds, s, A, sig64, pi, pj, npairs, cost); break;
        case 8: hipLaunchKernelGGL((k_dtw_pairs<T, 8>), grid, block, lds, s, A, sig64, pi, pj, npairs, cost); break;
        default: set_error("band too wide: 2w+1 must be <= 1024"); return DTWJ_EINVAL;
    }
    DTWJ_HIP(hipGetLastError());
    return 0;
}

}  // namespace

int launch_dtw_pairs(const JoinArgs &A, const double *sig64, const int64_t *pi, const int64_t *pj, int64_t npairs,
                     int32_t prec, double *cost, hipStream_t s) {
    if (npairs <= 0) return 0;
    if (prec == 64) return launch_pairs_t<double>(A, sig64, pi, pj, npairs, cost, s);
    return launch_pairs_t<float>(A, sig64, pi, pj, npairs, cost, s);
}

}  // namespace dtwj
