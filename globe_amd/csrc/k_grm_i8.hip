// Mixed-precision dense GRM product GZ = sig2g G S0 on the int8 matrix cores.
//
// The reference multiplies by the FP32 GRM in floating point (PCGsolver.cpp:1204-1212,
// varCompEst.cpp:336-347). Here the off-diagonal part of G is held as block fixed point:
// row i has the scale 2^e_i >= max_{j != i} |G_ij| (frexp exponent of that maximum, 0 for
// an all-zero row), and G_ij = 2^e_i sum_a d_a 2^(-6 - 8a) with four balanced int8
// digits d_0 .. d_3 (|d_0| <= 64, the rest in [-128, 127]) of rint(G_ij 2^(30 - e_i)).
// Those 30 bits resolve every entry at least as finely as its own FP32 rounding once
// the entry is within 2^-6 of the row maximum; the diagonal, far larger than the rest of
// its row, stays out of the digits and is applied in FP64 (G_ii S0_i). S0 gets the same
// form per column (scale 2^e_c >= max_j |S0_jc|). With L digits of each, the product
// keeps the digit pairs a, b < L with a + b <= L, grouped by weight class w = a + b,
// each class accumulated exactly in int32 over a chunk of at most 16384 columns of G
// (at most 4 pairs per class: |sum| <= 4 x 128 x 128 x 16384 = 2^30), then scaled by
// 2^(e_i + e_c - 12 - 8w) in FP64. The pairs left out (a + b > L) weigh at most about
// 2^(-8L - 4) of the scales each, below the truncation of each operand to L digits
// (2^(1 - 8L)); keeping only a + b < L would leave the class w = L, about 2^(4 - 8L).
// The PCG lowers L as its residual falls (inexact Krylov relaxation); L = 4 is within
// about 1e-9 of the FP64 product. The matrix cores do L(L+1)/2 + L - 1 int8 products
// per tile and column block, far from bounding the pass.
//
// Layout: digit tiles of the FULL square (both triangles, so no tile is read
// transposed): tile (I, J) at (I ntiles + J) x 16 KB, four 4 KB planes (digit a, then
// row-major 64 x 64 bytes), so a product with L digits streams the first L planes of
// every tile. Within a plane, 16-byte piece p of row r sits at slot p ^ f(r),
// f(r) = ((r >> 1) ^ (r >> 3)) & 3: the 16 rows that one MFMA operand read touches
// (one piece each) then cover every LDS bank once, with 32 or 64 banks. The S0 digits
// are [tile J][digit][column][64 rows] with the same swizzle by column.
//
// The product kernel: a workgroup owns the strip of 64 rows I and one chunk of its tile
// columns J, streams (tile I J, S0 slab J) stages global -> LDS (global_load_lds_dwordx4,
// 1 KB pieces) through a 3-slot ring, and each wave multiplies 16 rows x 16 columns with
// v_mfma_i32_16x16x64_i8 (one 64-wide K step per tile and digit pair). Blocks of 4 waves
// take 16 columns, of 8 waves 32 (two column groups share each tile). At 4 bytes per
// entry over the full square a pass reads 4 N^2 bytes, the same as the FP64 kernel's two
// reads of the FP32 triangle, but the matrix cores are no longer the bound.
#include <algorithm>
#include <cmath>

#include "globe_internal.h"

namespace globe {

namespace {

constexpr int kT = 64;                    // tile edge
constexpr int kPlaneB = kT * kT;          // bytes per digit plane of a tile
constexpr int kDig = 4;                   // digits held
constexpr int kTileDB = kDig * kPlaneB;   // bytes per digit tile
constexpr int kMaxChunkTiles = 256;       // tile columns per chunk partial (16384 columns)
constexpr int kI8Cols = 32;               // columns per pass
constexpr int kNS = 3;                    // ring slots

typedef int v4i __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

__device__ inline int64_t tile_index(int64_t I, int64_t J) { return I * (I + 1) / 2 + J; }  // J <= I
__device__ inline int swz_slot(int r) { return ((r >> 1) ^ (r >> 3)) & 3; }
// byte offset of (row r, column q) in a swizzled 64 x 64 plane
__device__ inline int swz(int r, int q) { return r * 64 + ((((q >> 4) ^ swz_slot(r)) & 3) << 4) + (q & 15); }

// Four balanced base-256 digits of v (|v| <= 2^30), most significant first.
__device__ inline void digits4(int v, int8_t d[4]) {
#pragma unroll
    for (int a = 3; a > 0; --a) {
        const int lo = (int)(int8_t)(v & 255);
        d[a] = (int8_t)lo;
        v = (v - lo) >> 8;
    }
    d[0] = (int8_t)v;
}

__device__ inline int exp_of(float m) {
    int e = 0;
    if (m > 0.f) (void)frexpf(m, &e);
    return e;
}

// rowmax[i] (float bits) = max over j != i of |G_ij|: one block per stored tile (I, J),
// J <= I; its row maxima go to rows 64 I + a, its column maxima to rows 64 J + b (G is
// symmetric). Non-negative floats order like their bit patterns.
__global__ __launch_bounds__(256) void k_grm_rowmax(const float* __restrict__ tiles, int32_t n,
                                                    uint32_t* __restrict__ rowmax) {
    const int I = blockIdx.y, J = blockIdx.x;
    if (J > I) return;
    __shared__ float sm[kT][kT + 1];
    const float* t = tiles + tile_index(I, J) * (kT * kT);
    for (int e = threadIdx.x; e < kT * kT; e += 256) {
        const int a = e / kT, b = e % kT;
        const int64_t i = (int64_t)I * kT + a, j = (int64_t)J * kT + b;
        sm[a][b] = (i == j) ? 0.f : fabsf(t[e]);
    }
    __syncthreads();
    const int x = threadIdx.x;
    if (x < kT) {
        float m = 0.f;
        for (int b = 0; b < kT; ++b) m = fmaxf(m, sm[x][b]);
        if ((int64_t)I * kT + x < n) atomicMax(rowmax + (int64_t)I * kT + x, __float_as_uint(m));
    } else if (x < 2 * kT) {
        const int b = x - kT;
        float m = 0.f;
        for (int a = 0; a < kT; ++a) m = fmaxf(m, sm[a][b]);
        if ((int64_t)J * kT + b < n) atomicMax(rowmax + (int64_t)J * kT + b, __float_as_uint(m));
    }
}

__global__ __launch_bounds__(256) void k_grm_rexp(const uint32_t* __restrict__ rowmax, int64_t npad,
                                                  int32_t* __restrict__ rexp) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < npad; i += (int64_t)gridDim.x * 256)
        rexp[i] = exp_of(__uint_as_float(rowmax[i]));
}

// Digit tile (I, J) of the full square: thread t takes row t / 4, columns 16 (t % 4) ..
// + 15 (from stored tile (I, J), or (J, I) transposed above the diagonal).
__global__ __launch_bounds__(256) void k_grm_digits(const float* __restrict__ tiles, int32_t ntiles,
                                                    const int32_t* __restrict__ rexp,
                                                    int8_t* __restrict__ G8) {
    const int J = blockIdx.x, I = blockIdx.y;
    const int r = threadIdx.x >> 2, p = threadIdx.x & 3;
    const int64_t i = (int64_t)I * kT + r;
    const int e = rexp[i];
    const float* src = J <= I ? tiles + tile_index(I, J) * (kT * kT) : tiles + tile_index(J, I) * (kT * kT);
    alignas(16) int8_t d[kDig][16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int b = 16 * p + q;
        const int64_t j = (int64_t)J * kT + b;
        float g = J <= I ? src[r * kT + b] : src[b * kT + r];
        if (i == j) g = 0.f;
        const int v = (int)rint(ldexp((double)g, 30 - e));
        int8_t dd[4];
        digits4(v, dd);
#pragma unroll
        for (int a = 0; a < kDig; ++a) d[a][q] = dd[a];
    }
    int8_t* dst = G8 + ((int64_t)I * ntiles + J) * kTileDB;
#pragma unroll
    for (int a = 0; a < kDig; ++a)
        *reinterpret_cast<uint4*>(dst + a * kPlaneB + swz(r, 16 * p)) = *reinterpret_cast<const uint4*>(d[a]);
}

// cmax[c] = max_j |S0[c][j]| as its bit pattern (non-negative doubles order like their
// bits), from a grid of (parts, columns) blocks; cmax starts at zero.
__global__ __launch_bounds__(256) void k_grm_colmax(const double* __restrict__ S0, int64_t ldn, int32_t n,
                                                    unsigned long long* __restrict__ cmax, const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    double m = 0.0;
    for (int64_t j = blockIdx.x * 256ll + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256)
        m = fmax(m, fabs(S0[c * ldn + j]));
    __shared__ double sh[256];
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && sh[0] > 0.0) atomicMax(cmax + c, (unsigned long long)__double_as_longlong(sh[0]));
}

// frexp exponent of a column maximum held as bits (0 for a zero column)
__device__ inline int col_exp(unsigned long long bits) {
    int e = 0;
    const double m = __longlong_as_double((long long)bits);
    if (m > 0.0) (void)frexp(m, &e);
    return e;
}

// S0 digits Sd[J][a][c][swizzled row] for rows j < 64 ntiles, columns c < cols (zero
// past n and kk); one thread per (row, column).
__global__ __launch_bounds__(256) void k_grm_s0_digits(const double* __restrict__ S0, int64_t ldn,
                                                       int32_t n, int kk, int cols, int32_t ntiles,
                                                       const unsigned long long* __restrict__ cmax,
                                                       int8_t* __restrict__ Sd, const int* skip) {
    if (skip && *skip) return;
    const int64_t total = (int64_t)ntiles * kT * cols;
    for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t j = e % ((int64_t)ntiles * kT);
        const int c = (int)(e / ((int64_t)ntiles * kT));
        int v = 0;
        if (c < kk && j < n) v = (int)rint(ldexp(S0[c * ldn + j], 30 - col_exp(cmax[c])));
        int8_t d[4];
        digits4(v, d);
        const int64_t J = j / kT;
        const int q = (int)(j % kT);
        const int off = ((((q >> 4) ^ swz_slot(c & 15)) & 3) << 4) + (q & 15);
#pragma unroll
        for (int a = 0; a < kDig; ++a) Sd[((J * kDig + a) * cols + c) * kT + off] = d[a];
    }
}

template <int CG, int L>
struct I8Cfg {
    static constexpr int kCols = 16 * CG;
    static constexpr int kWaves = 4 * CG;
    static constexpr int kTileKB = L * kPlaneB / 1024;      // 4 L
    static constexpr int kSlabKB = L * kCols * kT / 1024;   // L CG
    static constexpr int kCopies = kTileKB + kSlabKB;
    static constexpr int kPerWave = (kCopies + kWaves - 1) / kWaves;
    static constexpr int kStage = kCopies * 1024;
    static_assert((kNS - 2) * kPerWave < 64, "vmcnt field");
};

// One pass of up to 16 CG columns: part[h][c][i] = 2^(e_i + e_c) x the digit product of
// strip I over the tile columns of chunk h.
template <int CG, int L>
__global__ __launch_bounds__(256 * CG) void k_grm_i8(const int8_t* __restrict__ G8,
                                                     const int8_t* __restrict__ Sd,
                                                     const int32_t* __restrict__ rexp,
                                                     const unsigned long long* __restrict__ cmax, int kk,
                                                     int32_t n,
                                                     int32_t ntiles, int H, double* __restrict__ part,
                                                     const int* skip) {
    if (skip && *skip) return;
    using C = I8Cfg<CG, L>;
    __shared__ __attribute__((aligned(1024))) int8_t lds[kNS * C::kStage];
    const int I = blockIdx.x, h = blockIdx.y;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int rg = w & 3, cg = w >> 2;
    const int J0 = (int)((int64_t)h * ntiles / H), J1 = (int)((int64_t)(h + 1) * ntiles / H);
    const int nt = J1 - J0;
    v4i acc[L + 1];  // weight classes w = a + b = 0 .. L
#pragma unroll
    for (int q = 0; q <= L; ++q) acc[q] = v4i{0, 0, 0, 0};
    if (nt > 0) {
        // stage t: tile (I, J0 + t) and slab J0 + t (the last one repeated past the end,
        // into slots never read); copy p of a stage is 1 KB, waves take p = w, w + waves,
        // ... and repeat the last copy to even the count (same bytes, same place)
        auto issue = [&](int t, int slot) {
            const int64_t J = J0 + (t < nt ? t : nt - 1);
            const int8_t* tsrc = G8 + ((int64_t)I * ntiles + J) * kTileDB;
            const int8_t* ssrc = Sd + J * kDig * C::kCols * kT;
            int8_t* dst = lds + slot * C::kStage;
#pragma unroll
            for (int q = 0; q < C::kPerWave; ++q) {
                int p = w + q * C::kWaves;
                if (p >= C::kCopies) p = C::kCopies - 1;
                const int8_t* src = p < C::kTileKB ? tsrc + p * 1024 : ssrc + (p - C::kTileKB) * 1024;
                __builtin_amdgcn_global_load_lds((const void*)(src + lane * 16), (lds_void*)(dst + p * 1024), 16,
                                                 0, 0);
            }
        };
#pragma unroll
        for (int u = 0; u < kNS - 1; ++u) issue(u, u);
        const int r = 16 * rg + (lane & 15), piece = lane >> 4;
        const int cc = 16 * cg + (lane & 15);
        const int aoff = swz(r, 16 * piece);
        const int boff = cc * kT + ((((piece) ^ swz_slot(cc & 15)) & 3) << 4);
        int cs = 0;
        for (int t = 0; t < nt; ++t) {
            // this wave's copies of stage t have landed (those of the kNS - 2 later stages
            // may be in flight); after the barrier every wave's have, and every wave is done
            // with the slot the next copies overwrite
            asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"((kNS - 2) * C::kPerWave)
                         : "memory");
            issue(t + kNS - 1, cs == 0 ? kNS - 1 : cs - 1);
            __builtin_amdgcn_sched_barrier(0);
            const int8_t* tb = lds + cs * C::kStage;
            const int8_t* sb = tb + C::kTileKB * 1024;
            v4i A[L], B[L];
#pragma unroll
            for (int a = 0; a < L; ++a) A[a] = *reinterpret_cast<const v4i*>(tb + a * kPlaneB + aoff);
#pragma unroll
            for (int b = 0; b < L; ++b) B[b] = *reinterpret_cast<const v4i*>(sb + b * C::kCols * kT + boff);
#pragma unroll
            for (int a = 0; a < L; ++a)
#pragma unroll
                for (int b = 0; b < L && a + b <= L; ++b)
                    acc[a + b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[a], B[b], acc[a + b], 0, 0, 0);
            cs = cs + 1 == kNS ? 0 : cs + 1;
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    // D element v of lane l: (row 4 (l / 16) + v, column l % 16)
    const int col = 16 * cg + (lane & 15);
    if (col >= kk) return;
    const int ec = col_exp(cmax[col]);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int64_t i = (int64_t)I * kT + 16 * rg + 4 * (lane >> 4) + v;
        if (i >= n) continue;
        const int ei = rexp[i];
        double val = 0.0;
#pragma unroll
        for (int q = 0; q <= L; ++q) val = val + ldexp((double)acc[q][v], ei + ec - 12 - 8 * q);
        part[((int64_t)h * kI8Cols + col) * n + i] = val;
    }
}

// GZ[c][i] = sig2g_c (sum over chunks of the partials, in order, + G_ii S0[c][i])
__global__ __launch_bounds__(256) void k_grm_i8_finish(const double* __restrict__ part, int32_t n, int H,
                                                       const double* __restrict__ th,
                                                       const double* __restrict__ gdiag,
                                                       const double* __restrict__ S0, int64_t ldn,
                                                       double* __restrict__ GZ, int64_t ldg, const int* skip) {
    // no fused multiply-add: the product G_ii S0_i is rounded before the sum, so the
    // result is a fixed sequence of IEEE operations
#pragma clang fp contract(off)
    if (skip && *skip) return;
    const int c = blockIdx.y;
    const double sg = th[c * 5 + 1];
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        double v = 0.0;
        for (int h = 0; h < H; ++h) v = v + part[((int64_t)h * kI8Cols + c) * n + i];
        const double d = gdiag[i] * S0[c * ldn + i];
        v = v + d;
        GZ[c * ldg + i] = sg * v;
    }
}

template <int CG, int L>
void launch_i8(hipStream_t s, dim3 g, const int8_t* G8, const int8_t* Sd, const int32_t* rexp,
               const unsigned long long* cmax, int kk, int32_t n, int32_t ntiles, int H, double* part,
               const int* skip) {
    k_grm_i8<CG, L><<<g, 256 * CG, 0, s>>>(G8, Sd, rexp, cmax, kk, n, ntiles, H, part, skip);
}

int chunks_for(int32_t ntiles) { return std::max(4, (ntiles + kMaxChunkTiles - 1) / kMaxChunkTiles); }

size_t align1k(size_t b) { return (b + 1023) / 1024 * 1024; }

}  // namespace

size_t grm_digit_bytes(int32_t n) {
    const size_t T = (size_t)(n + kT - 1) / kT;
    return T * T * kTileDB;
}

size_t grm_digit_aux_bytes(int32_t n) {
    const size_t T = (size_t)(n + kT - 1) / kT;
    return 2 * T * kT * sizeof(int32_t);
}

void launch_grm_digits(hipStream_t s, const float* tiles, int32_t n, void* aux, int8_t* G8) {
    if (n <= 0) return;
    const int32_t ntiles = (n + kT - 1) / kT;
    const int64_t npad = (int64_t)ntiles * kT;
    uint32_t* rowmax = static_cast<uint32_t*>(aux);
    int32_t* rexp = reinterpret_cast<int32_t*>(rowmax + npad);
    HIP_CHECK(hipMemsetAsync(rowmax, 0, sizeof(uint32_t) * npad, s));
    k_grm_rowmax<<<dim3(ntiles, ntiles), 256, 0, s>>>(tiles, n, rowmax);
    HIP_CHECK(hipGetLastError());
    k_grm_rexp<<<(unsigned)std::min<int64_t>((npad + 255) / 256, 1024), 256, 0, s>>>(rowmax, npad, rexp);
    HIP_CHECK(hipGetLastError());
    k_grm_digits<<<dim3(ntiles, ntiles), 256, 0, s>>>(tiles, ntiles, rexp, G8);
    HIP_CHECK(hipGetLastError());
}

const int32_t* grm_digit_rexp(const void* aux, int32_t n) {
    const int64_t npad = (int64_t)((n + kT - 1) / kT) * kT;
    return reinterpret_cast<const int32_t*>(static_cast<const uint32_t*>(aux) + npad);
}

size_t grm_i8_part_bytes(int32_t n) {
    const int32_t ntiles = (n + kT - 1) / kT;
    const size_t H = (size_t)chunks_for(ntiles);
    return align1k(sizeof(double) * H * kI8Cols * (size_t)n) + align1k((size_t)ntiles * kDig * kI8Cols * kT) +
           1024;
}

void launch_grm_product_i8(hipStream_t s, const int8_t* G8, const int32_t* rexp, int32_t n, const double* S0,
                           int64_t ldn, int k, int L, const double* gdiag, const double* theta_cols, double* GZ,
                           int64_t ldg, const int* skip, void* scratch) {
    if (n == 0) return;
    L = std::min(std::max(L, 2), kDig);
    const int32_t ntiles = (n + kT - 1) / kT;
    const int H = chunks_for(ntiles);
    double* part = static_cast<double*>(scratch);
    int8_t* Sd = static_cast<int8_t*>(scratch) + align1k(sizeof(double) * H * kI8Cols * (size_t)n);
    unsigned long long* cmax =
        reinterpret_cast<unsigned long long*>(Sd + align1k((size_t)ntiles * kDig * kI8Cols * kT));
    const unsigned gm = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 2047) / 2048, 64));
    const dim3 g(ntiles, H);
    const unsigned gf = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
    for (int c = 0; c < k; c += kI8Cols) {
        const int kk = std::min(kI8Cols, k - c);
        const int CG = kk > 16 ? 2 : 1;
        const int cols = 16 * CG;
        launch_fill(s, reinterpret_cast<double*>(cmax), kk, 0.0);  // all-zero bits
        k_grm_colmax<<<dim3(gm, kk), 256, 0, s>>>(S0 + c * ldn, ldn, n, cmax, skip);
        HIP_CHECK(hipGetLastError());
        const unsigned gd = (unsigned)std::min<int64_t>(((int64_t)ntiles * kT * cols + 255) / 256, 4096);
        k_grm_s0_digits<<<gd, 256, 0, s>>>(S0 + c * ldn, ldn, n, kk, cols, ntiles, cmax, Sd, skip);
        HIP_CHECK(hipGetLastError());
        if (CG == 1) {
            if (L == 2) launch_i8<1, 2>(s, g, G8, Sd, rexp, cmax, kk, n, ntiles, H, part, skip);
            else if (L == 3) launch_i8<1, 3>(s, g, G8, Sd, rexp, cmax, kk, n, ntiles, H, part, skip);
            else launch_i8<1, 4>(s, g, G8, Sd, rexp, cmax, kk, n, ntiles, H, part, skip);
        } else {
            if (L == 2) launch_i8<2, 2>(s, g, G8, Sd, rexp, cmax, kk, n, ntiles, H, part, skip);
            else if (L == 3) launch_i8<2, 3>(s, g, G8, Sd, rexp, cmax, kk, n, ntiles, H, part, skip);
            else launch_i8<2, 4>(s, g, G8, Sd, rexp, cmax, kk, n, ntiles, H, part, skip);
        }
        HIP_CHECK(hipGetLastError());
        k_grm_i8_finish<<<dim3(gf, kk), 256, 0, s>>>(part, n, H, theta_cols + c * 5, gdiag, S0 + c * ldn, ldn,
                                                     GZ + c * ldg, ldg, skip);
        HIP_CHECK(hipGetLastError());
    }
}

}  // namespace globe
