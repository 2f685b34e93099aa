// Packed-genotype contraction on the int8 matrix cores.
//
//   Out[r][c] = sum_kappa dosage(r, kappa) * Z[c][kappa]
//
// The dosages 0/1/2 are exact int8. Each column of Z is written in block fixed
// point, q = round(z * 2^sh_c) with |q| < 2^(8L-2) (sh_c from the column's max |z|),
// and q is split into L balanced base-256 int8 digits. Every MFMA accumulates
// exactly in int32 (|acc| <= 256 K < 2^31 for K < 8.3M genotypes), so the product
// is exact for the quantized input; the only rounding is the quantization
// (2^-(8L-2) of the column max) and the final FP64 combination of the L digit sums.
//   L = 7: 54-bit fixed point, as accurate as an FP64 dot product (GLOBE_PREC_FP64)
//   L = 4: 30-bit fixed point (GLOBE_PREC_MIXED); 4 int8 digits cost the MFMA time of
//          two BF16 pieces and keep 30 bits with exact accumulation.
// Centering and scaling (w = (g - 2f) inv) are applied afterwards as the FP64
// rank-one correction sum_i w_is z_i = inv_s (P_s - 2 f_s T), T = sum_i z_i of the
// quantized z, so the operator is applied exactly to the quantized vector.
//
// Two kernels share the K order (a super-step is 256 genotypes = 64 packed bytes of
// a row; a lane decodes one packed dword = 16 genotypes per MFMA K-step):
//   thin (digit columns <= 16, the 1-RHS solve): v_mfma_i32_16x16x64_i8, 8 row
//     fragments of 16 rows per wave (512 rows per block), the digit fragments read
//     straight from global memory (L1 / L2), no LDS and no barrier;
//     A lane l: row l % 16, packed bytes 64 ss + 16 (l / 16) + 4 s for K-step s < 4.
//   wide (digit columns 17..128): v_mfma_i32_32x32x32_i8, 2 row fragments of 32 rows
//     and up to 4 column fragments of 32 per wave (256 rows per block), the digit
//     tile of the next super-step staged in a second LDS buffer while the current
//     one is consumed (one barrier per super-step);
//     A lane l: row l % 32, packed bytes 64 ss + 32 (l / 32) + 4 s for K-step s < 8.
// Decode: two masks and four v_perm_b32 byte-table lookups turn a dword of 16 codes
// into 16 int8 dosages in the order {0,4,8,12, 1,5,9,13, 2,6,10,14, 3,7,11,15}; the
// digit matrix Bq holds each lane's 16 digits in exactly that order.
// Output fragments (probed on gfx950): 16x16: element v of lane l is (row 4 (l/16) + v,
// column l % 16); 32x32: (row 8 (v/4) + 4 (l/32) + v % 4, column l % 32).
//
// Split-K is a power of two S; block b works on K-split b % S of row block b / S for
// S <= 8, and for S > 8 every XCD (b % 8) walks its splits one after another
// (split = b % 8 + 8 ((b / 8) / R)), so the blocks resident on one XCD share one
// K-chunk of Bq in that XCD's L2. Partial sums per split go to out[split][row][J16]
// and are added exactly (int64) in the epilogue.
#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "context.h"

namespace globe {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// dosage byte table for codes 0..3: 00 -> 2, 01 (missing) -> 0, 10 -> 1, 11 -> 0
constexpr uint32_t kDosTable = 0x00010002u;
constexpr uint32_t kZeroCodes = 0xFFFFFFFFu;  // 16 codes 11: dosage 0

__device__ inline v4i decode16(uint32_t w) {
    // byte j of (w >> 2d) & 0x03030303 is the code of genotype 4 j + d
    v4i d;
    d[0] = (int)__builtin_amdgcn_perm(0u, kDosTable, w & 0x03030303u);
    d[1] = (int)__builtin_amdgcn_perm(0u, kDosTable, (w >> 2) & 0x03030303u);
    d[2] = (int)__builtin_amdgcn_perm(0u, kDosTable, (w >> 4) & 0x03030303u);
    d[3] = (int)__builtin_amdgcn_perm(0u, kDosTable, (w >> 6) & 0x03030303u);
    return d;
}

__device__ inline uint32_t dword_of(const uint4& v, int s) {
    return s == 0 ? v.x : s == 1 ? v.y : s == 2 ? v.z : v.w;
}

// block b -> (row block, split)
__device__ inline void block_coords(int64_t b, int64_t R, int S, int64_t& rb, int& sp) {
    if (S <= 8) {
        sp = (int)(b % S);
        rb = b / S;
    } else {
        const int64_t t = b / 8;
        sp = (int)(b % 8) + 8 * (int)(t / R);
        rb = t % R;
    }
}

// ---------------------------------------------------------------- thin kernel
// RT row fragments of 16 rows per wave (64 RT rows per block). Row blocks are read
// without bounds checks: the packed layouts are padded to a multiple of 512 rows.
// Two register sets alternate so the next super-step's loads overlap the current
// one's decode and MFMAs.
template <int RT>
__device__ inline void thin_load(const uint8_t* base, int64_t frag_stride, uint32_t loff, int64_t ss,
                                 const uint4* bq, uint4 (&A)[RT], uint4 (&B)[4]) {
#pragma unroll
    for (int a = 0; a < RT; ++a)
        A[a] = *reinterpret_cast<const uint4*>(base + a * frag_stride + loff + ss * 64);
#pragma unroll
    for (int s = 0; s < 4; ++s) B[s] = bq[(4 * ss + s) * 64];
}

template <int RT>
__device__ inline void thin_mma(const uint4 (&A)[RT], const uint4 (&B)[4], v4i (&acc)[RT]) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const v4i bv = *reinterpret_cast<const v4i*>(&B[s]);
#pragma unroll
        for (int a = 0; a < RT; ++a)
            acc[a] = __builtin_amdgcn_mfma_i32_16x16x64_i8(decode16(dword_of(A[a], s)), bv, acc[a], 0, 0, 0);
    }
}

template <int RT>
__global__ __launch_bounds__(256) void k_geno_thin(const uint8_t* __restrict__ rows, int64_t pitch,
                                                   int64_t nrows, int64_t n_ss, int64_t ss_per_split,
                                                   int64_t R, int S, const uint8_t* __restrict__ Bq,
                                                   int32_t* __restrict__ out, const int* skip) {
    if (skip && *skip) return;
    int64_t rb;
    int sp;
    block_coords(blockIdx.x, R, S, rb, sp);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, rr = lane & 15;
    const int64_t row_base = rb * (64 * RT) + wave * (16 * RT);
    const uint8_t* base = rows + row_base * pitch;  // uniform per wave
    const int64_t fstride = 16 * pitch;
    const uint32_t loff = (uint32_t)(rr * pitch + 16 * g);
    const int64_t ss0 = (int64_t)sp * ss_per_split;
    const int64_t ss1 = ss0 + ss_per_split < n_ss ? ss0 + ss_per_split : n_ss;
    const uint4* bq = reinterpret_cast<const uint4*>(Bq) + lane;  // [ks][64 lanes] x 16 B

    v4i acc[RT];
#pragma unroll
    for (int a = 0; a < RT; ++a) acc[a] = v4i{0, 0, 0, 0};
    uint4 A0[RT], A1[RT], B0[4], B1[4];
    if (ss0 < ss1) thin_load<RT>(base, fstride, loff, ss0, bq, A0, B0);
    for (int64_t ss = ss0; ss < ss1; ss += 2) {
        if (ss + 1 < ss1) thin_load<RT>(base, fstride, loff, ss + 1, bq, A1, B1);
        thin_mma<RT>(A0, B0, acc);
        if (ss + 1 >= ss1) break;
        if (ss + 2 < ss1) thin_load<RT>(base, fstride, loff, ss + 2, bq, A0, B0);
        thin_mma<RT>(A1, B1, acc);
    }
    int32_t* o = out + (int64_t)sp * nrows * 16;
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t r = row_base + a * 16 + 4 * g + v;
            if (r < nrows) o[r * 16 + rr] = acc[a][v];
        }
}

// ---------------------------------------------------------------- wide kernel
constexpr int kWideRows = 4 * 64;  // 4 waves x 2 fragments of 32 rows

template <int JT>
__global__ __launch_bounds__(256) void k_geno_wide(const uint8_t* __restrict__ rows, int64_t pitch,
                                                   int64_t nrows, int64_t n_ss, int64_t ss_per_split,
                                                   int64_t R, int S, const uint8_t* __restrict__ Bq,
                                                   int32_t* __restrict__ out, const int* skip) {
    if (skip && *skip) return;
    // two buffers of one super-step of digits: 8 K-steps x JT fragments x 1 KB
    __shared__ __attribute__((aligned(16))) uint8_t lds[2][8 * JT * 1024];
    int64_t rb;
    int sp;
    block_coords(blockIdx.x, R, S, rb, sp);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 5, rr = lane & 31;
    const int64_t row_base = rb * kWideRows + wave * 64;
    const uint8_t* base = rows + row_base * pitch;  // uniform per wave
    const int64_t fstride = 32 * pitch;
    const uint32_t loff = (uint32_t)(rr * pitch + 32 * g);
    const int64_t ss0 = (int64_t)sp * ss_per_split;
    const int64_t ss1 = ss0 + ss_per_split < n_ss ? ss0 + ss_per_split : n_ss;

    v16i acc[2][JT];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < JT; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0;
    // digits of one super-step: 8 JT KB = 512 JT uint4, thread t copies t + 256 q, q < 2 JT
    constexpr int kCopies = 2 * JT;
    const uint4* bsrc = reinterpret_cast<const uint4*>(Bq) + threadIdx.x;
    uint4 an[2][2], bn[kCopies];
    if (ss0 < ss1) {
        uint4* d4 = reinterpret_cast<uint4*>(lds[0]) + threadIdx.x;
#pragma unroll
        for (int q = 0; q < kCopies; ++q) d4[256 * q] = bsrc[ss0 * (512 * JT) + 256 * q];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            an[a][0] = *reinterpret_cast<const uint4*>(base + a * fstride + loff + ss0 * 64);
            an[a][1] = *reinterpret_cast<const uint4*>(base + a * fstride + loff + ss0 * 64 + 16);
        }
    }
    __syncthreads();
    int cur = 0;
    for (int64_t ss = ss0; ss < ss1; ++ss) {
        const uint4 c00 = an[0][0], c01 = an[0][1], c10 = an[1][0], c11 = an[1][1];
        // the whole next super-step (packed rows and digits) is requested up front; on
        // the last super-step the current one is re-read instead (in bounds, unused), so
        // the prefetch registers are written unconditionally and stay in VGPRs
        const int64_t nx = ss + 1 < ss1 ? ss + 1 : ss;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            an[a][0] = *reinterpret_cast<const uint4*>(base + a * fstride + loff + nx * 64);
            an[a][1] = *reinterpret_cast<const uint4*>(base + a * fstride + loff + nx * 64 + 16);
        }
#pragma unroll
        for (int q = 0; q < kCopies; ++q) bn[q] = bsrc[nx * (512 * JT) + 256 * q];
        const uint8_t* cl = lds[cur] + lane * 16;
        // digit fragments of K-step s + 1 are read from LDS while step s multiplies
        v4i bcur[JT];
#pragma unroll
        for (int b = 0; b < JT; ++b) bcur[b] = *reinterpret_cast<const v4i*>(cl + b * 1024);
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            v4i bnx[JT];
            if (s < 7)
#pragma unroll
                for (int b = 0; b < JT; ++b)
                    bnx[b] = *reinterpret_cast<const v4i*>(cl + ((s + 1) * JT + b) * 1024);
            const v4i a0 = decode16(dword_of(s < 4 ? c00 : c01, s & 3));
            const v4i a1 = decode16(dword_of(s < 4 ? c10 : c11, s & 3));
#pragma unroll
            for (int b = 0; b < JT; ++b) {
                acc[0][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, bcur[b], acc[0][b], 0, 0, 0);
                acc[1][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, bcur[b], acc[1][b], 0, 0, 0);
            }
            if (s < 7)
#pragma unroll
                for (int b = 0; b < JT; ++b) bcur[b] = bnx[b];
        }
        {
            uint4* d4 = reinterpret_cast<uint4*>(lds[cur ^ 1]) + threadIdx.x;
#pragma unroll
            for (int q = 0; q < kCopies; ++q) d4[256 * q] = bn[q];
        }
        __syncthreads();
        cur ^= 1;
    }
    constexpr int J16 = 32 * JT;
    int32_t* o = out + (int64_t)sp * nrows * J16;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int64_t r = row_base + a * 32 + 8 * (v >> 2) + 4 * g + (v & 3);
            if (r < nrows)
#pragma unroll
                for (int b = 0; b < JT; ++b) o[r * J16 + b * 32 + rr] = acc[a][b][v];
        }
}

// ---------------------------------------------------------------- reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    return v;
}
__device__ inline double block_sum(double v) {
    __shared__ double sh[4];
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = sh[0] + sh[1] + sh[2] + sh[3];
    return r;
}
__device__ inline double block_max(double v) {
    __shared__ double sh[4];
    v = wave_max(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    return r;
}

// per-column max |v| partials
__global__ __launch_bounds__(256) void k_colmax_parts(int64_t rows, const double* __restrict__ v,
                                                      int64_t ldv, double* part, const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    double m = 0.0;
    for (int64_t q = blockIdx.x * 256ll + threadIdx.x; q < rows; q += (int64_t)gridDim.x * 256)
        m = fmax(m, fabs(v[c * ldv + q]));
    const double r = block_max(m);
    if (threadIdx.x == 0) part[c * kMaxParts + blockIdx.x] = r;
}

// One wave per column: max (want_max) or ordered sum of the column's partials; the
// scale exponent sh_c = (8L - 2) - e_c with max|v| < 2^e_c when sh != nullptr.
__global__ __launch_bounds__(256) void k_finish_cols(const double* __restrict__ part, int nparts,
                                                     int k, int want_max, int limbs, int* sh,
                                                     double* out, const int* skip) {
    if (skip && *skip) return;
    const int lane = threadIdx.x & 63;
    for (int c = threadIdx.x >> 6; c < k; c += 4) {
        double v = 0.0;
        for (int j = lane; j < nparts; j += 64) {
            const double x = part[c * kMaxParts + j];
            v = want_max ? fmax(v, x) : v + x;
        }
        v = want_max ? wave_max(v) : wave_sum(v);
        if (lane == 0) {
            if (out) out[c] = v;
            if (sh) {
                int e = 0;
                if (v > 0.0) (void)frexp(v, &e);  // v < 2^e
                sh[c] = (v > 0.0 && isfinite(v)) ? (8 * limbs - 2) - e : 0;
            }
        }
    }
}

// Quantize v (k columns, K rows valid, Kpad total) into the digit layout of the thin
// (wide = 0) or wide (wide = 1) kernel and emit per-block partials of
// sum_kappa w_kappa q_kappa 2^-sh (T for pass 1 with w = 1, C for pass 2 with w = 2 f).
__global__ __launch_bounds__(256) void k_quantize(const double* __restrict__ v, int64_t ldv,
                                                  int64_t K, int64_t Kpad, int k, int limbs,
                                                  int nfrag, int wide, const int* __restrict__ sh,
                                                  const double* __restrict__ wts,
                                                  uint8_t* __restrict__ Bq, double* part,
                                                  const int* skip) {
    if (skip && *skip) return;
    const int j = blockIdx.y;  // digit column
    const int c = j / limbs, l = j % limbs;
    const bool live = c < k;
    const int fw = wide ? 32 : 16;
    const int jf = j / fw, jj = j % fw;
    double acc = 0.0;
    for (int64_t kap = blockIdx.x * 256ll + threadIdx.x; kap < Kpad; kap += (int64_t)gridDim.x * 256) {
        int dig = 0;
        if (live && kap < K) {
            const double z = v[c * ldv + kap];
            const long long q = llrint(ldexp(z, sh[c]));
            if (l == 0) acc += (wts ? wts[kap] : 1.0) * ldexp((double)q, -sh[c]);
            long long t = q;
            for (int d = 0; d <= l; ++d) {
                int b = (int)(t & 0xFF);
                if (b >= 128) b -= 256;
                dig = b;
                t = (t - b) >> 8;
            }
        }
        const int64_t ss = kap >> 8;
        const int rem = (int)(kap & 255);
        const int w = rem & 15;
        const int qslot = 4 * (w & 3) + (w >> 2);
        int64_t ks;
        int ln;
        if (wide) {
            ks = 8 * ss + ((rem >> 4) & 7);
            ln = 32 * (rem >> 7) + jj;
        } else {
            ks = 4 * ss + ((rem >> 4) & 3);
            ln = 16 * (rem >> 6) + jj;
        }
        Bq[((ks * nfrag + jf) * 64 + ln) * 16 + qslot] = (uint8_t)(int8_t)dig;
    }
    if (live && l == 0) {
        const double r = block_sum(acc);
        if (threadIdx.x == 0) part[c * kMaxParts + blockIdx.x] = r;
    }
}

// Epilogues: a block takes 64 rows x k columns. Phase 1 reads the digit sums with
// consecutive threads on consecutive digits (coalesced), phase 2 writes the FP64
// results column-major with consecutive threads on consecutive rows.
constexpr int kEpiRows = 64;

__device__ inline double combine(const int32_t* __restrict__ out, int nsplit, int64_t split_stride,
                                 int64_t base, int limbs) {
    double v = 0.0;
    for (int l = 0; l < limbs; ++l) {
        long long s = 0;
        for (int sp = 0; sp < nsplit; ++sp) s += out[sp * split_stride + base + l];
        v += ldexp((double)s, 8 * l);
    }
    return v;
}

// pass 1 (mode 0): u = sig2g / M * inv^2 * (P - 2 f T), per-block colmax of u
// pass 2 (mode 1): GZ = P - C
__global__ __launch_bounds__(256) void k_epilogue(const int32_t* __restrict__ out, int nsplit,
                                                  int64_t rows, int J16, int limbs, int k,
                                                  const int* __restrict__ sh, int mode,
                                                  const double* __restrict__ TC,
                                                  const double* __restrict__ twof,
                                                  const double* __restrict__ inv,
                                                  const double* __restrict__ th, double inv_m,
                                                  double* __restrict__ res, int64_t ldr,
                                                  double* part, const int* skip) {
    if (skip && *skip) return;
    __shared__ double tile[kEpiRows][kMaxCols + 1];
    __shared__ double cmax[kMaxCols];
    const int64_t split_stride = rows * J16;
    const int64_t ntiles = (rows + kEpiRows - 1) / kEpiRows;
    for (int c = threadIdx.x; c < k; c += 256) cmax[c] = 0.0;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t r0 = t * kEpiRows;
        __syncthreads();
        for (int q = threadIdx.x; q < kEpiRows * k; q += 256) {
            const int rl = q / k, c = q % k;
            const int64_t r = r0 + rl;
            double val = 0.0;
            if (r < rows) {
                const double P =
                    ldexp(combine(out, nsplit, split_stride, r * J16 + c * limbs, limbs), -sh[c]);
                if (mode == 0) {
                    const double iv = inv[r];
                    val = th[c * 5 + 1] * inv_m * iv * iv * (P - twof[r] * TC[c]);
                } else {
                    val = P - TC[c];
                }
            }
            tile[rl][c] = val;
        }
        __syncthreads();
        for (int q = threadIdx.x; q < kEpiRows * k; q += 256) {
            const int c = q / kEpiRows, rl = q % kEpiRows;
            const int64_t r = r0 + rl;
            if (r < rows) res[c * ldr + r] = tile[rl][c];
        }
        if (mode == 0) {
            // running per-column maxima (one wave per column, fixed order)
            for (int c = threadIdx.x >> 6; c < k; c += 4) {
                double m = 0.0;
                const int rl = threadIdx.x & 63;
                if (r0 + rl < rows) m = fabs(tile[rl][c]);
                m = wave_max(m);
                if ((threadIdx.x & 63) == 0) cmax[c] = fmax(cmax[c], m);
            }
        }
    }
    __syncthreads();
    if (mode == 0 && part)
        for (int c = threadIdx.x; c < k; c += 256) part[c * kMaxParts + blockIdx.x] = cmax[c];
}

// Split-K: enough blocks to fill the chip (>= 4 per CU), each split >= 8 super-steps;
// wide contractions split at least 8 ways so each XCD streams one K-chunk of Bq.
int choose_split(int64_t R, int64_t n_ss, bool wide) {
    int S = 1;
    while (R * S < 1024 && S < 64) S <<= 1;
    if (wide && S < 8) S = 8;
    while (S > 1 && n_ss / S < 8) S >>= 1;
    return S;
}

struct Variant {
    int rt;  // thin: row fragments per wave (4 or 8); wide: 2
    int S;   // split-K
};

int64_t rows_per_block(bool wide, int rt) { return wide ? kWideRows : 64 * rt; }

void launch_variant(hipStream_t s, bool wide, int JT, const Variant& v, const uint8_t* rows,
                    int64_t pitch, int64_t nrows, int64_t nss, const uint8_t* Bq, int32_t* out,
                    const int* skip) {
    const int64_t R = (nrows + rows_per_block(wide, v.rt) - 1) / rows_per_block(wide, v.rt);
    const int64_t per = (nss + v.S - 1) / v.S;
    const unsigned grid = (unsigned)(R * v.S);
    if (!wide) {
        if (v.rt == 8)
            k_geno_thin<8><<<grid, 256, 0, s>>>(rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip);
        else
            k_geno_thin<4><<<grid, 256, 0, s>>>(rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip);
    } else if (JT == 1) {
        k_geno_wide<1><<<grid, 256, 0, s>>>(rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip);
    } else if (JT == 2) {
        k_geno_wide<2><<<grid, 256, 0, s>>>(rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip);
    } else {
        k_geno_wide<4><<<grid, 256, 0, s>>>(rows, pitch, nrows, nss, per, R, v.S, Bq, out, skip);
    }
    HIP_CHECK(hipGetLastError());
}

}  // namespace

// Tile autotuner: the first contraction of each shape (pass, rows, K, digit columns)
// times every candidate variant once on the live data with HIP events and keeps the
// fastest; every variant computes the same exact integer sums, so the choice never
// changes a result. Candidates: thin rows-per-wave 4 / 8, split-K S/2, S, 2 S around
// choose_split().
static std::vector<Variant> candidates(bool wide, int64_t nrows, int64_t nss) {
    std::vector<Variant> v;
    const int rts[2] = {8, 4};
    for (int ri = 0; ri < (wide ? 1 : 2); ++ri) {
        const int rt = wide ? 2 : rts[ri];
        const int64_t R = (nrows + rows_per_block(wide, rt) - 1) / rows_per_block(wide, rt);
        const int S0 = choose_split(R, nss, wide);
        for (int S : {S0 / 2, S0, S0 * 2}) {
            if (S < 1 || S > 64 || (S > 1 && nss / S < 4)) continue;
            v.push_back({rt, S});
        }
    }
    return v;
}

// GZ (n x k) = sig2g G S0 for the packed-genotype cohort with `limbs` int8 digits.
void mfma_geno_product(globe_ctx* c, const double* S0, int64_t ldn, int k, double* GZ, int64_t ldg,
                       const int* skip, int limbs, const double* theta_cols) {
    hipStream_t s = c->stream;
    const int64_t n = c->n, m = c->m;
    const int J = k * limbs;
    require(J <= 128, GLOBE_E_INVALID, "k * limbs must be <= 128 for the MFMA contraction");
    const bool wide = J > 16;
    const int JT = wide ? (J <= 32 ? 1 : J <= 64 ? 2 : 4) : 1;
    const int J16 = wide ? 32 * JT : 16;
    const int nfrag = wide ? JT : 1;
    const int64_t Kpad1 = c->pitch_s * 4, Kpad2 = c->pitch_i * 4;  // genotypes per padded row
    const int64_t nss1 = Kpad1 / 256, nss2 = Kpad2 / 256;
    // the largest split any candidate may use bounds the partial-sum buffer
    const size_t bq1 = (size_t)Kpad1 * J16, bq2 = (size_t)Kpad2 * J16;
    int smax1 = 1, smax2 = 1;
    for (const Variant& v : candidates(wide, m, nss1)) smax1 = std::max(smax1, v.S);
    for (const Variant& v : candidates(wide, n, nss2)) smax2 = std::max(smax2, v.S);
    const size_t ob = std::max((size_t)smax1 * m, (size_t)smax2 * n) * J16 * 4;
    const size_t bq = std::max(bq1, bq2);
    c->w_mix.ensure(round_up(bq, 256) + round_up(ob, 256) + 2 * kMaxCols * sizeof(int));
    uint8_t* Bq = c->w_mix.as<uint8_t>();
    int32_t* out = reinterpret_cast<int32_t*>(Bq + round_up(bq, 256));
    int* sh1 = reinterpret_cast<int*>(Bq + round_up(bq, 256) + round_up(ob, 256));
    int* sh2 = sh1 + kMaxCols;
    c->w_U.ensure(sizeof(double) * (size_t)(m > 0 ? m : 1) * k);
    c->w_T.ensure(sizeof(double) * kMaxCols);
    c->w_C.ensure(sizeof(double) * kMaxCols);
    c->w_part2.ensure(sizeof(double) * kMaxParts * kMaxCols);
    double* part = c->w_part2.as<double>();
    double* U = c->w_U.as<double>();
    const double* th = theta_cols;
    const double flops = 2.0 * (double)n * (double)m * k;

    // run one pass with the tuned variant (tuning it first if this shape is new)
    auto contract = [&](int pass, const uint8_t* rows, int64_t pitch, int64_t nrows, int64_t K,
                        int64_t Kpad, int64_t nss) -> int {
        const std::string key = std::to_string(pass) + ":" + std::to_string(nrows) + ":" +
                                std::to_string(Kpad) + ":" + std::to_string(J16);
        auto it = c->tuned.find(key);
        if (it == c->tuned.end()) {
            const std::vector<Variant> cand = candidates(wide, nrows, nss);
            hipEvent_t e0, e1;
            HIP_CHECK(hipEventCreate(&e0));
            HIP_CHECK(hipEventCreate(&e1));
            float best = 1e30f;
            int bi = 0;
            for (size_t q = 0; q < cand.size(); ++q) {
                HIP_CHECK(hipEventRecord(e0, s));
                launch_variant(s, wide, JT, cand[q], rows, pitch, nrows, nss, Bq, out, nullptr);
                HIP_CHECK(hipEventRecord(e1, s));
                HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                if (std::getenv("GLOBE_TUNE_LOG"))
                    std::fprintf(stderr, "[globe tune] %s rows/wave-frags %d split %d: %.3f ms\n",
                                 key.c_str(), cand[q].rt, cand[q].S, ms);
                if (ms < best) {
                    best = ms;
                    bi = (int)q;
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            it = c->tuned.emplace(key, std::make_pair(cand[bi].rt, cand[bi].S)).first;
        }
        const Variant v{it->second.first, it->second.second};
        if (c->prof_on) {
            globe_ctx::ProfEvent e{};
            HIP_CHECK(hipEventCreate(&e.a));
            HIP_CHECK(hipEventCreate(&e.b));
            e.pass = pass;
            e.iter = c->prof_iter;
            e.bytes = (double)nrows * (double)((K + 3) / 4);
            e.flops = flops;
            e.issued = 2.0 * (double)nrows * (double)Kpad * J16;
            HIP_CHECK(hipEventRecord(e.a, s));
            launch_variant(s, wide, JT, v, rows, pitch, nrows, nss, Bq, out, skip);
            HIP_CHECK(hipEventRecord(e.b, s));
            c->prof.push_back(e);
        } else {
            launch_variant(s, wide, JT, v, rows, pitch, nrows, nss, Bq, out, skip);
        }
        return v.S;
    };

    // pass 1: P1[s][c] = sum_i g_is S0[c][i] on the SNP-major rows
    const int npn = parts_for(n);
    k_colmax_parts<<<dim3(npn, k), 256, 0, s>>>(n, S0, ldn, part, skip);
    k_finish_cols<<<1, 256, 0, s>>>(part, npn, k, 1, limbs, sh1, nullptr, skip);
    const int qg1 = parts_for(Kpad1);
    k_quantize<<<dim3(qg1, J16), 256, 0, s>>>(S0, ldn, n, Kpad1, k, limbs, nfrag, wide, sh1, nullptr,
                                              Bq, part, skip);
    k_finish_cols<<<1, 256, 0, s>>>(part, qg1, k, 0, limbs, nullptr, c->w_T.as<double>(), skip);
    HIP_CHECK(hipGetLastError());
    const int sp1 = contract(0, c->d_snp.as<uint8_t>(), c->pitch_s, m, n, Kpad1, nss1);
    const int64_t eb1 = std::min<int64_t>((m + kEpiRows - 1) / kEpiRows, kMaxParts);
    k_epilogue<<<(unsigned)eb1, 256, 0, s>>>(out, sp1, m, J16, limbs, k, sh1, 0, c->w_T.as<double>(),
                                             c->d_twof.as<double>(), c->d_inv.as<double>(), th,
                                             1.0 / (double)m, U, m, part, skip);
    HIP_CHECK(hipGetLastError());
    // pass 2: GZ[c][i] = sum_s g_is u[c][s] - C[c] on the individual-major rows
    k_finish_cols<<<1, 256, 0, s>>>(part, (int)eb1, k, 1, limbs, sh2, nullptr, skip);
    const int qg2 = parts_for(Kpad2);
    k_quantize<<<dim3(qg2, J16), 256, 0, s>>>(U, m, m, Kpad2, k, limbs, nfrag, wide, sh2,
                                              c->d_twof.as<double>(), Bq, part, skip);
    k_finish_cols<<<1, 256, 0, s>>>(part, qg2, k, 0, limbs, nullptr, c->w_C.as<double>(), skip);
    HIP_CHECK(hipGetLastError());
    const int sp2 = contract(1, c->d_ind.as<uint8_t>(), c->pitch_i, n, m, Kpad2, nss2);
    const int64_t eb2 = std::min<int64_t>((n + kEpiRows - 1) / kEpiRows, kMaxParts);
    k_epilogue<<<(unsigned)eb2, 256, 0, s>>>(out, sp2, n, J16, limbs, k, sh2, 1, c->w_C.as<double>(),
                                             nullptr, nullptr, nullptr, 0.0, GZ, ldg, nullptr, skip);
    HIP_CHECK(hipGetLastError());
}

}  // namespace globe
