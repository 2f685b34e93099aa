// Dense symmetric GRM product GZ = sig2g G S0 from the FP32 packed lower triangle
// (GCTA .grm.bin, element (i, j), j <= i, at i(i+1)/2 + j; PCGsolver.cpp:878-881,
// :1204-1212; varCompEst.cpp:336-347), FP64 accumulation, and the GRM builder that
// writes that triangle from the resident genotypes.
//
// Product: a workgroup owns a strip of 64 output rows and one of kGrmChunks ranges of
// 64-column tiles of the full symmetric row strip. Each 64 x 64 tile is staged in LDS
// with coalesced loads -- from the stored block (I, J) when J < I, transposed from the
// stored block (J, I) when J > I, mirrored when J == I -- next to the 64 x 16 slice of
// S0, and multiplied on the FP64 matrix cores (an FP64 VALU form of the same tile
// product is bound by LDS bandwidth: 2 operand reads per 4 FMAs); a group of up to 16
// columns takes one pass over the triangle. The strip partials of the chunks are added
// in chunk order, so the result is deterministic. Each stored element is read twice (once for its row strip, once
// for its column strip); the triangle is far larger than any cache.
#include <algorithm>

#include "globe_internal.h"

namespace globe {

namespace {

constexpr int kGT = 64;        // tile edge
constexpr int kGrmChunks = 4;  // column ranges per row strip

__device__ inline int64_t lt_index(int64_t i, int64_t j) { return i * (i + 1) / 2 + j; }

typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kGP = kGT + 4;  // LDS row pitch of a tile (floats)

// Thread t holds tile elements (a, b) = (t / 64 + 4 q, t % 64), q < 16, of tile (I, J):
// KIND 0 (J < I): G[i0 + a][j0 + b] from the stored block (I, J);
// KIND 1 (J > I): G[i0 + b][j0 + a] from the stored block (J, I) (row j0 + a, coalesced in b);
// KIND 2 (J == I): the diagonal block, mirrored.
template <int KIND>
__device__ inline void tile_load(const float* __restrict__ lt, int32_t n, int64_t i0, int64_t j0, int t,
                                 float (&v)[16]) {
    const int b = t & 63;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int a = (t >> 6) + 4 * q;
        int64_t idx;
        if (KIND == 0) {
            const int64_t i = i0 + a;
            idx = i < n ? lt_index(i, j0 + b) : 0;
        } else if (KIND == 1) {
            const int64_t j = j0 + a;
            idx = j < n ? lt_index(j, i0 + b) : 0;
        } else {
            const int64_t i = i0 + a, j = j0 + b;
            idx = (i < n && j < n) ? (b <= a ? lt_index(i, j) : lt_index(j, i)) : 0;
        }
        v[q] = lt[idx];
    }
}

template <int KIND>
__device__ inline void tile_store(float* __restrict__ g, int32_t n, int64_t i0, int64_t j0, int t,
                                  const float (&v)[16]) {
    const int b = t & 63;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int a = (t >> 6) + 4 * q;
        if (KIND == 0)
            g[a * kGP + b] = i0 + a < n ? v[q] : 0.0f;
        else if (KIND == 1)
            g[b * kGP + a] = j0 + a < n ? v[q] : 0.0f;
        else
            g[a * kGP + b] = (i0 + a < n && j0 + b < n) ? v[q] : 0.0f;
    }
}

constexpr int kGC = 16;       // columns of S0 per pass

// One pass of up to 16 columns: out (64 x 16) += G_tile (64 x 64) S0_slice (64 x 16) per
// tile on the FP64 matrix cores (v_mfma_f64_16x16x4f64): wave w owns rows 16 w .. 16 w
// + 15 of the strip; A lane l = G[16 w + l % 16][4 s + l / 16] (FP32 tile in LDS,
// widened), B lane l = S0[4 s + l / 16][l % 16], D element v of lane l = (row
// 4 v + l / 16, column l % 16) (probed on gfx950). Columns past kk are zero.
__global__ __launch_bounds__(256) void k_grm_tiles(const float* __restrict__ lt, int32_t n,
                                                   const double* __restrict__ S0, int64_t ldn, int kk,
                                                   int32_t ntiles, double* __restrict__ part,
                                                   const int* skip) {
    if (skip && *skip) return;
    __shared__ float g[kGT * kGP];
    __shared__ double sh[kGT * kGC];  // [k][c]
    const int I = blockIdx.x, h = blockIdx.y;
    const int64_t i0 = (int64_t)I * kGT;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int J0 = (int)((int64_t)h * ntiles / kGrmChunks), J1 = (int)((int64_t)(h + 1) * ntiles / kGrmChunks);
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    for (int J = J0; J < J1; ++J) {
        const int64_t j0 = (int64_t)J * kGT;
        // all 16 + 4 loads of the tile are issued before the first LDS store (indices
        // clamped in bounds; out-of-range elements become 0 when stored)
        float v[16];
        const int kind = J < I ? 0 : (J > I ? 1 : 2);
        if (kind == 0)
            tile_load<0>(lt, n, i0, j0, t, v);
        else if (kind == 1)
            tile_load<1>(lt, n, i0, j0, t, v);
        else
            tile_load<2>(lt, n, i0, j0, t, v);
        double sv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q, kq = e >> 4, c = e & 15;
            const int64_t j = j0 + kq;
            const bool good = c < kk && j < n;
            const double x = S0[good ? c * ldn + j : 0];
            sv[q] = good ? x : 0.0;
        }
        __syncthreads();  // the previous tile's MFMAs have read g and sh
        if (kind == 0)
            tile_store<0>(g, n, i0, j0, t, v);
        else if (kind == 1)
            tile_store<1>(g, n, i0, j0, t, v);
        else
            tile_store<2>(g, n, i0, j0, t, v);
#pragma unroll
        for (int q = 0; q < 4; ++q) sh[t + 256 * q] = sv[q];
        __syncthreads();
        const float* gr = g + (16 * w + (lane & 15)) * kGP + (lane >> 4);
        const double* sb = sh + (lane >> 4) * kGC + (lane & 15);
#pragma unroll 4
        for (int s4 = 0; s4 < kGT / 4; ++s4)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64((double)gr[4 * s4], sb[4 * s4 * kGC], acc, 0, 0, 0);
    }
    const int col = lane & 15;
    if (col >= kk) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int64_t i = i0 + 16 * w + 4 * v + (lane >> 4);
        if (i < n) part[((int64_t)h * kGC + col) * n + i] = acc[v];
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
    // one pass over the triangle per group of up to 16 columns
    for (int c = 0; c < k; c += kGC) {
        const int kk = std::min(kGC, k - c);
        k_grm_tiles<<<g, 256, 0, s>>>(lt, n, S0 + c * ldn, ldn, kk, ntiles, part, skip);
        HIP_CHECK(hipGetLastError());
        k_grm_finish<<<dim3(gf, kk), 256, 0, s>>>(part, n, kGC, theta_cols + c * 5, GZ + c * ldg, ldg,
                                                  skip);
        HIP_CHECK(hipGetLastError());
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
