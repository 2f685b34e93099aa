// Dense symmetric GRM product GZ = sig2g G S0 from the FP32 packed lower triangle
// (GCTA .grm.bin, element (i, j), j <= i, at i(i+1)/2 + j; PCGsolver.cpp:878-881,
// :1204-1212; varCompEst.cpp:336-347), FP64 accumulation, and the GRM builder that
// writes that triangle from the resident genotypes.
//
// Product: a workgroup owns a strip of 64 output rows and one of kGrmChunks ranges of
// 64-column tiles of the full symmetric row strip. Each 64 x 64 tile is staged in LDS
// with coalesced loads -- from the stored block (I, J) when J < I, transposed from the
// stored block (J, I) when J > I, mirrored when J == I -- next to the 64 x k slice of
// S0, and every thread accumulates a 2 x (k/8) (k >= 8) or 1 x 1 register tile of
// outputs. The strip partials of the chunks are added in chunk order, so the result
// is deterministic. Each stored element is read twice (once for its row strip, once
// for its column strip); the triangle is far larger than any cache.
#include <algorithm>

#include "globe_internal.h"

namespace globe {

namespace {

constexpr int kGT = 64;        // tile edge
constexpr int kGrmChunks = 4;  // column ranges per row strip

__device__ inline int64_t lt_index(int64_t i, int64_t j) { return i * (i + 1) / 2 + j; }

template <int KC>
__global__ __launch_bounds__(256) void k_grm_tiles(const float* __restrict__ lt, int32_t n,
                                                   const double* __restrict__ S0, int64_t ldn,
                                                   int32_t ntiles, double* __restrict__ part,
                                                   const int* skip) {
    if (skip && *skip) return;
    __shared__ float g[kGT][kGT + 1];
    __shared__ double s[KC][kGT];
    const int I = blockIdx.x, h = blockIdx.y;
    const int64_t i0 = (int64_t)I * kGT;
    const int t = threadIdx.x;
    const int J0 = (int)((int64_t)h * ntiles / kGrmChunks), J1 = (int)((int64_t)(h + 1) * ntiles / kGrmChunks);
    // register tile: KC >= 8: rows 2 rp, 2 rp + 1 and columns cg * CPT .. + CPT;
    // KC < 8: row t % 64, column t / 64 (threads past KC columns idle)
    constexpr bool kWide = KC >= 8;
    constexpr int RPT = kWide ? 2 : 1, CPT = kWide ? KC / 8 : 1;
    const int r0 = kWide ? 2 * (t & 31) : (t & 63);
    const int c0 = kWide ? (t >> 5) * CPT : (t >> 6);
    const bool live = c0 < KC;
    double acc[RPT][CPT];
#pragma unroll
    for (int a = 0; a < RPT; ++a)
#pragma unroll
        for (int b = 0; b < CPT; ++b) acc[a][b] = 0.0;
    for (int J = J0; J < J1; ++J) {
        const int64_t j0 = (int64_t)J * kGT;
        __syncthreads();
        if (J < I) {
            for (int e = t; e < kGT * kGT; e += 256) {
                const int rr = e >> 6, cc = e & 63;
                const int64_t i = i0 + rr, j = j0 + cc;
                g[rr][cc] = i < n ? lt[lt_index(i, j)] : 0.0f;
            }
        } else if (J > I) {
            for (int e = t; e < kGT * kGT; e += 256) {
                const int jj = e >> 6, ii = e & 63;
                const int64_t i = i0 + ii, j = j0 + jj;
                g[ii][jj] = j < n ? lt[lt_index(j, i)] : 0.0f;
            }
        } else {
            for (int e = t; e < kGT * kGT; e += 256) {
                const int rr = e >> 6, cc = e & 63;
                const int64_t i = i0 + rr, j = j0 + cc;
                float v = 0.0f;
                if (i < n && j < n) v = cc <= rr ? lt[lt_index(i, j)] : lt[lt_index(j, i)];
                g[rr][cc] = v;
            }
        }
        for (int e = t; e < KC * kGT; e += 256) {
            const int c = e >> 6, jj = e & 63;
            const int64_t j = j0 + jj;
            s[c][jj] = j < n ? S0[c * ldn + j] : 0.0;
        }
        __syncthreads();
        if (live) {
#pragma unroll 8
            for (int jj = 0; jj < kGT; ++jj) {
                double gv[RPT], sv[CPT];
#pragma unroll
                for (int a = 0; a < RPT; ++a) gv[a] = (double)g[r0 + a][jj];
#pragma unroll
                for (int b = 0; b < CPT; ++b) sv[b] = s[c0 + b][jj];
#pragma unroll
                for (int a = 0; a < RPT; ++a)
#pragma unroll
                    for (int b = 0; b < CPT; ++b) acc[a][b] = fma(gv[a], sv[b], acc[a][b]);
            }
        }
    }
    if (!live) return;
#pragma unroll
    for (int a = 0; a < RPT; ++a) {
        const int64_t i = i0 + r0 + a;
        if (i < n)
#pragma unroll
            for (int b = 0; b < CPT; ++b) part[((int64_t)h * KC + c0 + b) * n + i] = acc[a][b];
    }
}

// GZ[c][i] = sig2g_c * sum over chunks of the strip partials (fixed order)
__global__ __launch_bounds__(256) void k_grm_finish(const double* __restrict__ part, int32_t n, int kc,
                                                    const double* __restrict__ th,
                                                    double* __restrict__ GZ, int64_t ldg,
                                                    const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    const double sg = th[c * 5 + 1];
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        double v = 0.0;
#pragma unroll
        for (int h = 0; h < kGrmChunks; ++h) v += part[((int64_t)h * kc + c) * n + i];
        GZ[c * ldg + i] = sg * v;
    }
}

// lt[i(i+1)/2 + j] = (float) G[i][j] for the columns j0 .. j0 + k - 1 of G given in
// Gc (n x k, column-major), rows i >= j
__global__ __launch_bounds__(256) void k_grm_scatter(const double* __restrict__ Gc, int64_t ldg,
                                                     int32_t n, int64_t j0, int k,
                                                     float* __restrict__ lt) {
    const int c = blockIdx.y;
    const int64_t j = j0 + c;
    for (int64_t i = j + blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        lt[lt_index(i, j)] = (float)Gc[c * ldg + i];
}

// S0 = columns j0 .. j0 + k - 1 of the n x n identity
__global__ __launch_bounds__(256) void k_eye_cols(double* __restrict__ S0, int64_t ld, int32_t n,
                                                  int64_t j0, int k) {
    const int c = blockIdx.y;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        S0[c * ld + i] = i == j0 + c ? 1.0 : 0.0;
}

__global__ __launch_bounds__(256) void k_grm_diag(const float* __restrict__ lt, int32_t n,
                                                  double* __restrict__ d) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        d[i] = (double)lt[lt_index(i, i)];
}

}  // namespace

void launch_eye_cols(hipStream_t s, double* S0, int64_t ld, int32_t n, int64_t j0, int k) {
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1024));
    k_eye_cols<<<dim3(gx, k), 256, 0, s>>>(S0, ld, n, j0, k);
    HIP_CHECK(hipGetLastError());
}

void launch_grm_diag(hipStream_t s, const float* lt, int32_t n, double* d) {
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1024));
    k_grm_diag<<<gx, 256, 0, s>>>(lt, n, d);
    HIP_CHECK(hipGetLastError());
}

void launch_grm_product(hipStream_t s, const float* lt, int32_t n, const double* S0, int64_t ldn,
                        int k, const double* theta_cols, double* GZ, int64_t ldg, const int* skip,
                        double* part) {
    if (n == 0) return;
    const int32_t ntiles = (n + kGT - 1) / kGT;
    const dim3 g(ntiles, kGrmChunks);
    const unsigned gf = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
    int c = 0;
    while (c < k) {
        const int left = k - c;
        const double* sc = S0 + c * ldn;
        double* gc = GZ + c * ldg;
        const double* tc = theta_cols + c * 5;
        int kc;
        if (left >= 16) {
            kc = 16;
            k_grm_tiles<16><<<g, 256, 0, s>>>(lt, n, sc, ldn, ntiles, part, skip);
        } else if (left >= 8) {
            kc = 8;
            k_grm_tiles<8><<<g, 256, 0, s>>>(lt, n, sc, ldn, ntiles, part, skip);
        } else if (left >= 4) {
            kc = 4;
            k_grm_tiles<4><<<g, 256, 0, s>>>(lt, n, sc, ldn, ntiles, part, skip);
        } else if (left >= 2) {
            kc = 2;
            k_grm_tiles<2><<<g, 256, 0, s>>>(lt, n, sc, ldn, ntiles, part, skip);
        } else {
            kc = 1;
            k_grm_tiles<1><<<g, 256, 0, s>>>(lt, n, sc, ldn, ntiles, part, skip);
        }
        HIP_CHECK(hipGetLastError());
        k_grm_finish<<<dim3(gf, kc), 256, 0, s>>>(part, n, kc, tc, gc, ldg, skip);
        HIP_CHECK(hipGetLastError());
        c += kc;
    }
}

size_t grm_product_part_bytes(int32_t n) { return sizeof(double) * (size_t)kGrmChunks * 16 * (size_t)n; }

void launch_grm_scatter(hipStream_t s, const double* Gc, int64_t ldg, int32_t n, int64_t j0, int k,
                        float* lt) {
    if (k <= 0) return;
    const int64_t rows = n - j0;
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((rows + 255) / 256, 1024));
    k_grm_scatter<<<dim3(gx, k), 256, 0, s>>>(Gc, ldg, n, j0, k, lt);
    HIP_CHECK(hipGetLastError());
}

}  // namespace globe
