// Dense symmetric GRM product GZ = sig2g G S0 from the FP32 packed lower triangle
// (GCTA .grm.bin, element (i, j), j <= i, at i(i+1)/2 + j; PCGsolver.cpp:878-881,
// :1204-1212; varCompEst.cpp:336-347). FP64 accumulation.
//
// One thread per output row i: the stored row part j <= i is a per-thread stream,
// the column part j > i (element (j, i)) is coalesced across the threads of a block
// and S0[j] is a broadcast there.
#include "globe_internal.h"

namespace globe {

namespace {

template <int KC>
__global__ __launch_bounds__(kBlock) void k_grm_rows(const float* __restrict__ lt, int32_t n,
                                                     const double* __restrict__ S0, int64_t ldn,
                                                     const double* __restrict__ th,
                                                     double* __restrict__ GZ, int64_t ldg,
                                                     const int* skip) {
    if (skip && *skip) return;
    const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (i >= n) return;
    double acc[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[c] = 0.0;
    const float* row = lt + i * (i + 1) / 2;
    for (int64_t j = 0; j <= i; ++j) {
        const double g = (double)row[j];
#pragma unroll
        for (int c = 0; c < KC; ++c) acc[c] = fma(g, S0[c * ldn + j], acc[c]);
    }
    for (int64_t j = i + 1; j < n; ++j) {
        const double g = (double)lt[j * (j + 1) / 2 + i];
#pragma unroll
        for (int c = 0; c < KC; ++c) acc[c] = fma(g, S0[c * ldn + j], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < KC; ++c) GZ[c * ldg + i] = th[c * 5 + 1] * acc[c];
}

}  // namespace

void launch_grm_product(hipStream_t s, const float* lt, int32_t n, const double* S0, int64_t ldn,
                        int k, const double* theta_cols, double* GZ, int64_t ldg, const int* skip) {
    if (n == 0) return;
    const unsigned g = (unsigned)((n + kBlock - 1) / kBlock);
    int c = 0;
    while (c < k) {
        const int left = k - c;
        const double* sc = S0 + c * ldn;
        double* gc = GZ + c * ldg;
        const double* tc = theta_cols + c * 5;
        if (left >= 8) {
            k_grm_rows<8><<<g, kBlock, 0, s>>>(lt, n, sc, ldn, tc, gc, ldg, skip);
            c += 8;
        } else if (left >= 4) {
            k_grm_rows<4><<<g, kBlock, 0, s>>>(lt, n, sc, ldn, tc, gc, ldg, skip);
            c += 4;
        } else if (left >= 2) {
            k_grm_rows<2><<<g, kBlock, 0, s>>>(lt, n, sc, ldn, tc, gc, ldg, skip);
            c += 2;
        } else {
            k_grm_rows<1><<<g, kBlock, 0, s>>>(lt, n, sc, ldn, tc, gc, ldg, skip);
            c += 1;
        }
        HIP_CHECK(hipGetLastError());
    }
}

}  // namespace globe
