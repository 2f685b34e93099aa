// Packed-genotype kernels.
//
// HBM layout (both resident for the life of the cohort):
//   snp-major  : m kept SNP rows x pitch_s bytes, individual i of SNP s in bits
//                2(i%4) of byte i/4 (PLINK 1 SNP-major, PCGsolver.cpp:825-827);
//   indiv-major: n individual rows x pitch_i bytes, SNP s in bits 2(s%4) of byte s/4.
// Pitches are multiples of 128 bytes (512 genotypes).
// The resident layouts hold each genotype as its 2-bit dosage (0, 1, 2), not the
// PLINK code: 00 -> 2, 10 -> 1, 11 -> 0, and a missing call 01 -> 0 (PCGsolver.cpp:
// 83-94, missing -> 0 at :837). After QC a missing call and code 11 enter every
// product identically, so the conversion loses nothing, and a dosage field is
// turned into an int8 byte with one mask. Phantom genotypes of the last byte and
// all padding hold dosage 0 (zero bytes), so padding never contributes.
#include "globe_internal.h"

namespace globe {

namespace {

// one packed byte of PLINK codes -> the same four genotypes as 2-bit dosages
__device__ inline uint32_t codes_to_dosages(uint32_t b) {
    const uint32_t L = b & 0x55u, H = (b >> 1) & 0x55u;
    const uint32_t one = H & ~L;            // code 10
    const uint32_t two = ~H & ~L & 0x55u;   // code 00
    return one | (two << 1);
}

// ------------------------------------------------------------------ QC
// Per SNP: A1 allele count 2*#00 + #10 and missing count #01 over the first n
// individuals (PCGsolver.cpp:711-724; excluding phantom pairs is what
// subtractPhantomGenotypes :590-613 achieves), integer exact; then the keep rule of
// :726-732 in the reference's FP32 order with IEEE round-to-nearest conversions and
// divisions: maf = (float)alt / (2.0f * (float)n_obs), missRate = (float)miss /
// (float)N, keep iff maf >= minMAF && maf <= 1 - minMAF && missRate <= maxMiss.
__global__ __launch_bounds__(kBlock) void k_bed_qc(const uint8_t* __restrict__ rows, int64_t m,
                                                   int64_t pitch, int32_t n, float min_maf,
                                                   float max_miss, int64_t* __restrict__ alt,
                                                   int64_t* __restrict__ miss, float* __restrict__ maf,
                                                   uint8_t* __restrict__ keep) {
    __shared__ int sh_a[kBlock / 64], sh_m[kBlock / 64];
    const int64_t nbytes = (n + 3) / 4;
    for (int64_t s = blockIdx.x; s < m; s += gridDim.x) {
        const uint8_t* row = rows + s * pitch;
        int a = 0, mi = 0;
        for (int64_t j = threadIdx.x; j < nbytes; j += kBlock) {
            const uint32_t b = row[j];
            const int64_t valid = n - 4 * j;
            const uint32_t mask = valid >= 4 ? 0x55u : (0x55u & ((1u << (2 * valid)) - 1u));
            const uint32_t L = b & 0x55u, H = (b >> 1) & 0x55u;
            a += 2 * __popc(~L & ~H & mask) + __popc(H & ~L & mask);
            mi += __popc(L & ~H & mask);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_down(a, o, 64);
            mi += __shfl_down(mi, o, 64);
        }
        __syncthreads();
        if ((threadIdx.x & 63) == 0) {
            sh_a[threadIdx.x >> 6] = a;
            sh_m[threadIdx.x >> 6] = mi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t ta = 0, tm = 0;
            for (int w = 0; w < kBlock / 64; ++w) {
                ta += sh_a[w];
                tm += sh_m[w];
            }
            alt[s] = ta;
            miss[s] = tm;
            if (maf) {
                const int n_obs = n - (int)tm;
                const float f = n_obs > 0 ? __fdiv_rn(__ll2float_rn(ta), __fmul_rn(2.0f, __int2float_rn(n_obs)))
                                          : 0.0f;
                const float mr = __fdiv_rn(__ll2float_rn(tm), __int2float_rn(n));
                maf[s] = f;
                keep[s] = (f >= min_maf && f <= __fsub_rn(1.0f, min_maf) && mr <= max_miss) ? 1 : 0;
            }
        }
    }
}

// Destination rows of one chunk's kept rows, in SNP order: pos[q] = *base + (kept
// rows before q) for a kept row q, else -1; then *base += the chunk's kept count.
// One block: each thread counts kScanPer consecutive flags, the block scans the
// thread counts (wave shuffles, then across waves in LDS), tile after tile.
constexpr int kScanThreads = 1024, kScanPer = 8;
__global__ __launch_bounds__(kScanThreads) void k_keep_scan(const uint8_t* __restrict__ keep, int64_t cnt,
                                                            int64_t* __restrict__ base,
                                                            int64_t* __restrict__ pos) {
    __shared__ int64_t wsum[kScanThreads / 64];
    __shared__ int64_t carry;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) carry = *base;
    __syncthreads();
    constexpr int64_t kTile = (int64_t)kScanThreads * kScanPer;
    for (int64_t t0 = 0; t0 < cnt; t0 += kTile) {
        const int64_t q0 = t0 + (int64_t)t * kScanPer;
        int c = 0;
#pragma unroll
        for (int e = 0; e < kScanPer; ++e) c += (q0 + e < cnt && keep[q0 + e]) ? 1 : 0;
        // inclusive scan of c over the wave
        int v = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(v, o, 64);
            if (lane >= o) v += u;
        }
        if (lane == 63) wsum[w] = v;
        __syncthreads();
        int64_t before = carry;
        for (int q = 0; q < w; ++q) before += wsum[q];
        before += v - c;  // exclusive within the wave
        int64_t r = before;
#pragma unroll
        for (int e = 0; e < kScanPer; ++e) {
            const int64_t q = q0 + e;
            if (q < cnt) {
                const bool k = keep[q] != 0;
                pos[q] = k ? r : -1;
                r += k ? 1 : 0;
            }
        }
        __syncthreads();  // every thread has read carry and wsum
        if (t == kScanThreads - 1) carry = r;
        __syncthreads();
    }
    if (t == 0) *base = carry;
}

// Copy the rows with pos[q] >= 0 to row pos[q] of dst as 2-bit dosages, with phantom
// genotypes of the last byte and the pitch padding set to dosage 0.
__global__ __launch_bounds__(kBlock) void k_compact_pos(const uint8_t* __restrict__ src,
                                                        int64_t pitch_src,
                                                        const int64_t* __restrict__ pos, int64_t cnt,
                                                        int32_t n, uint8_t* __restrict__ dst,
                                                        int64_t pitch_dst) {
    const int64_t nbytes = (n + 3) / 4;
    const int rem = n & 3;
    const uint32_t live = rem == 0 ? 0xFFu : (1u << (2 * rem)) - 1u;  // fields of the last byte
    for (int64_t q = blockIdx.x; q < cnt; q += gridDim.x) {
        const int64_t r = pos[q];
        if (r < 0) continue;
        const uint8_t* s = src + q * pitch_src;
        uint8_t* d = dst + r * pitch_dst;
        for (int64_t j = threadIdx.x; j < pitch_dst; j += kBlock) {
            uint32_t v = 0;
            if (j < nbytes) {
                v = codes_to_dosages(s[j]);
                if (j == nbytes - 1) v &= live;
            }
            d[j] = (uint8_t)v;
        }
    }
}

// SNP-major -> individual-major. One block per tile of 256 SNPs x 64 individuals:
// rows of 16 bytes are staged in LDS, each thread emits 16 output bytes (64 SNPs)
// of one individual.
__global__ __launch_bounds__(kBlock) void k_transpose(const uint8_t* __restrict__ snp, int64_t m,
                                                      int64_t pitch_s, int32_t n,
                                                      uint8_t* __restrict__ ind, int64_t pitch_i) {
    __shared__ uint8_t tile[256][17];
    const int64_t s0 = (int64_t)blockIdx.x * 256;
    const int64_t i0 = (int64_t)blockIdx.y * 64;
    const int t = threadIdx.x;
    {
        const int64_t s = s0 + t;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (s < m) v = *reinterpret_cast<const uint4*>(snp + s * pitch_s + i0 / 4);
        const uint8_t* vb = reinterpret_cast<const uint8_t*>(&v);
#pragma unroll
        for (int q = 0; q < 16; ++q) tile[t][q] = vb[q];
    }
    __syncthreads();
    const int il = t >> 2;         // individual within tile
    const int part = t & 3;        // which 64-SNP group
    const int64_t i = i0 + il;
    if (i >= n) return;
    uint8_t out[16];
#pragma unroll
    for (int ob = 0; ob < 16; ++ob) {
        uint32_t byte = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int sl = part * 64 + ob * 4 + q;
            const uint32_t dose = (tile[sl][il >> 2] >> (2 * (il & 3))) & 3u;
            byte |= dose << (2 * q);
        }
        out[ob] = (uint8_t)byte;
    }
    *reinterpret_cast<uint4*>(ind + i * pitch_i + s0 / 4 + part * 16) =
        *reinterpret_cast<const uint4*>(out);
}

// SplitMix64 finalizer (Steele, Lea & Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014) as a counter-based hash.
__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Synthetic SNP-major rows: f_s ~ U(f_lo, f_hi); per genotype two Bernoulli(f_s)
// draws (Binomial(2, f)) and a Bernoulli(miss) missing flag, from one 64-bit hash
// split into three 21-bit uniforms.
__global__ __launch_bounds__(kBlock) void k_synth_bed(uint8_t* __restrict__ rows, int64_t m,
                                                      int64_t pitch, int32_t n, uint64_t seed,
                                                      float f_lo, float f_hi, float miss) {
    const int64_t nbytes = (n + 3) / 4;
    const double inv21 = 1.0 / 2097152.0;
    for (int64_t s = blockIdx.y; s < m; s += gridDim.y) {
        const uint64_t hs = mix64(seed ^ (0xA5A5A5A5ull + (uint64_t)s * 0x100000001B3ull));
        const double f = f_lo + (f_hi - f_lo) * ((double)(hs >> 11) * (1.0 / 9007199254740992.0));
        uint8_t* row = rows + s * pitch;
        for (int64_t j = blockIdx.x * (int64_t)kBlock + threadIdx.x; j < pitch;
             j += (int64_t)gridDim.x * kBlock) {
            uint32_t byte = 0xFF;
            if (j < nbytes) {
                byte = 0;
                for (int q = 0; q < 4; ++q) {
                    const int64_t i = 4 * j + q;
                    uint32_t code = 3u;
                    if (i < n) {
                        const uint64_t h = mix64(hs + (uint64_t)i * 0x9E3779B97F4A7C15ull);
                        const double u1 = (double)(h & 0x1FFFFF) * inv21;
                        const double u2 = (double)((h >> 21) & 0x1FFFFF) * inv21;
                        const double u3 = (double)((h >> 42) & 0x1FFFFF) * inv21;
                        const int d = (u1 < f) + (u2 < f);
                        code = d == 2 ? 0u : (d == 1 ? 2u : 3u);
                        if (u3 < miss) code = 1u;
                    }
                    byte |= code << (2 * q);
                }
            }
            row[j] = (uint8_t)byte;
        }
    }
}

// twof = 2 f, inv = 1/sqrt(2 f (1 - f)) or 1 when 2 f (1 - f) <= 1e-12 (FP64 from the
// stored FP32 frequency; the reference evaluates the same formula in FP32,
// PCGsolver.cpp:1159-1161).
__global__ __launch_bounds__(kBlock) void k_snp_consts(const float* __restrict__ freq, int64_t m,
                                                       double* __restrict__ twof,
                                                       double* __restrict__ inv) {
    for (int64_t s = blockIdx.x * (int64_t)kBlock + threadIdx.x; s < m;
         s += (int64_t)gridDim.x * kBlock) {
        const double f = (double)freq[s];
        const double d = 2.0 * f * (1.0 - f);
        twof[s] = 2.0 * f;
        inv[s] = d > 1e-12 ? 1.0 / sqrt(d) : 1.0;
    }
}

// diag(G)_i = (1/m) sum_s ((g_is - 2 f_s) inv_s)^2 (PCGsolver.cpp:959-987) from the
// individual-major rows; the three possible values per SNP are staged in LDS.
__global__ __launch_bounds__(kBlock) void k_gdiag(const uint8_t* __restrict__ ind, int64_t pitch_i,
                                                  int32_t n, int64_t m,
                                                  const double* __restrict__ twof,
                                                  const double* __restrict__ inv,
                                                  double* __restrict__ gdiag) {
    __shared__ double cv[3][kGenoTile];
    const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    const uint8_t* row = ind + i * pitch_i;
    double acc = 0.0;
    for (int64_t k0 = 0; k0 < m; k0 += kGenoTile) {
        __syncthreads();
        for (int q = threadIdx.x; q < kGenoTile; q += kBlock) {
            const int64_t s = k0 + q;
            double c0 = 0.0, c1 = 0.0, c2 = 0.0;
            if (s < m) {
                const double a = twof[s], w = inv[s] * inv[s];
                c0 = a * a * w;
                c1 = (1.0 - a) * (1.0 - a) * w;
                c2 = (2.0 - a) * (2.0 - a) * w;
            }
            cv[0][q] = c0;
            cv[1][q] = c1;
            cv[2][q] = c2;
        }
        __syncthreads();
        if (i < n) {
            const uint4* q4 = reinterpret_cast<const uint4*>(row + k0 / 4);
#pragma unroll 2
            for (int j = 0; j < kGenoTile / 64; ++j) {
                const uint4 w = q4[j];
                const uint32_t wd[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int d = 0; d < 4; ++d)
#pragma unroll
                    for (int g = 0; g < 16; ++g) {
                        const int kk = j * 64 + d * 16 + g;
                        acc += cv[(wd[d] >> (2 * g)) & 3u][kk];
                    }
            }
        }
    }
    if (i < n) gdiag[i] = m > 0 ? acc / (double)m : 0.0;
}

// out[c][r] = sum_k dosage(r, k) v[c][k] in FP64: one thread per packed row, the
// K-tile of v staged in LDS and read as a broadcast by the whole block.
template <int KC>
__global__ __launch_bounds__(kBlock) void k_rowdot(const uint8_t* __restrict__ rows, int64_t pitch,
                                                   int64_t nrows, int64_t nk,
                                                   const double* __restrict__ v, int64_t ldv,
                                                   double* __restrict__ out, int64_t ldo,
                                                   const int* skip) {
    if (skip && *skip) return;
    __shared__ double sv[KC][kGenoTile];
    const int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    const uint8_t* row = rows + r * pitch;
    double acc[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[c] = 0.0;
    for (int64_t k0 = 0; k0 < nk; k0 += kGenoTile) {
        __syncthreads();
        for (int q = threadIdx.x; q < KC * kGenoTile; q += kBlock) {
            const int c = q / kGenoTile, kk = q % kGenoTile;
            const int64_t kg = k0 + kk;
            sv[c][kk] = kg < nk ? v[c * ldv + kg] : 0.0;
        }
        __syncthreads();
        if (r < nrows) {
            const uint4* q4 = reinterpret_cast<const uint4*>(row + k0 / 4);
            for (int j = 0; j < kGenoTile / 64; ++j) {
                const uint4 w = q4[j];
                const uint32_t wd[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int d = 0; d < 4; ++d)
#pragma unroll
                    for (int g = 0; g < 16; ++g) {
                        const int kk = j * 64 + d * 16 + g;
                        const double dv = (double)((wd[d] >> (2 * g)) & 3u);
#pragma unroll
                        for (int c = 0; c < KC; ++c) acc[c] = fma(dv, sv[c][kk], acc[c]);
                    }
            }
        }
    }
    if (r < nrows)
#pragma unroll
        for (int c = 0; c < KC; ++c) out[c * ldo + r] = acc[c];
}

__device__ inline double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

__device__ inline double block_sum_d(double v) {
    __shared__ double sh[kBlock / 64];
    v = wave_sum_d(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) r += sh[i];
    return r;
}

__global__ __launch_bounds__(kBlock) void k_colsum_parts(int64_t rows, const double* __restrict__ v,
                                                         int64_t ldv, double* part,
                                                         const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    double acc = 0.0;
    for (int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x; q < rows;
         q += (int64_t)gridDim.x * kBlock)
        acc += v[c * ldv + q];
    const double s = block_sum_d(acc);
    if (threadIdx.x == 0) part[c * kMaxParts + blockIdx.x] = s;
}

// Pass-1 epilogue (centering as a rank-one correction of the uncentred product):
//   sum_i w_is z_i = inv_s (P1_s - 2 f_s T), T = sum_i z_i;
//   u_s = (sig2g / M) inv_s^2 (P1_s - 2 f_s T)  so that  GZ_i = sum_s g_is u_s - C,
//   C = sum_s 2 f_s u_s (partials here).
__global__ __launch_bounds__(kBlock) void k_center_snp(int64_t m, const double* __restrict__ P1,
                                                       int64_t ld1, const double* __restrict__ T,
                                                       const double* __restrict__ twof,
                                                       const double* __restrict__ inv,
                                                       const double* __restrict__ th,
                                                       double inv_m, double* __restrict__ u,
                                                       int64_t ldu, double* part_c,
                                                       const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    const double scale = th[c * 5 + 1] * inv_m, tc = T[c];
    double acc = 0.0;
    for (int64_t s = blockIdx.x * (int64_t)kBlock + threadIdx.x; s < m;
         s += (int64_t)gridDim.x * kBlock) {
        const double iv = inv[s];
        const double uv = scale * iv * iv * (P1[c * ld1 + s] - twof[s] * tc);
        u[c * ldu + s] = uv;
        acc += twof[s] * uv;
    }
    const double sum = block_sum_d(acc);
    if (threadIdx.x == 0) part_c[c * kMaxParts + blockIdx.x] = sum;
}

__global__ __launch_bounds__(kBlock) void k_center_ind(int32_t n, const double* __restrict__ P2,
                                                       int64_t ld2, const double* __restrict__ Cs,
                                                       double* __restrict__ GZ, int64_t ldg,
                                                       const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    const double cc = Cs[c];
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        GZ[c * ldg + i] = P2[c * ld2 + i] - cc;
}

}  // namespace

#define LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

void launch_bed_qc(hipStream_t s, const uint8_t* rows, int64_t m, int64_t pitch, int32_t n,
                   float min_maf, float max_miss, int64_t* alt, int64_t* miss, float* maf,
                   uint8_t* keep) {
    int64_t g = m < 65536 ? m : 65536;
    if (g < 1) return;
    k_bed_qc<<<(unsigned)g, kBlock, 0, s>>>(rows, m, pitch, n, min_maf, max_miss, alt, miss, maf, keep);
    LAUNCH_CHECK();
}

void launch_keep_scan(hipStream_t s, const uint8_t* keep, int64_t cnt, int64_t* base, int64_t* pos) {
    if (cnt < 1) return;
    k_keep_scan<<<1, kScanThreads, 0, s>>>(keep, cnt, base, pos);
    LAUNCH_CHECK();
}

void launch_compact_pos(hipStream_t s, const uint8_t* src, int64_t pitch_src, const int64_t* pos,
                        int64_t cnt, int32_t n, uint8_t* dst, int64_t pitch_dst) {
    int64_t g = cnt < 65536 ? cnt : 65536;
    if (g < 1) return;
    k_compact_pos<<<(unsigned)g, kBlock, 0, s>>>(src, pitch_src, pos, cnt, n, dst, pitch_dst);
    LAUNCH_CHECK();
}

void launch_transpose_geno(hipStream_t s, const uint8_t* snp, int64_t m, int64_t pitch_s, int32_t n,
                           uint8_t* ind, int64_t pitch_i) {
    if (m == 0 || n == 0) return;
    dim3 g((unsigned)((m + 255) / 256), (unsigned)((n + 63) / 64));
    k_transpose<<<g, kBlock, 0, s>>>(snp, m, pitch_s, n, ind, pitch_i);
    LAUNCH_CHECK();
}

void launch_synth_bed(hipStream_t s, uint8_t* rows, int64_t m, int64_t pitch, int32_t n,
                      uint64_t seed, float f_lo, float f_hi, float miss) {
    if (m == 0) return;
    unsigned gx = (unsigned)((pitch + kBlock - 1) / kBlock);
    unsigned gy = (unsigned)(m < 65535 ? m : 65535);
    k_synth_bed<<<dim3(gx, gy), kBlock, 0, s>>>(rows, m, pitch, n, seed, f_lo, f_hi, miss);
    LAUNCH_CHECK();
}

void launch_snp_consts(hipStream_t s, const float* freq, int64_t m, double* twof, double* inv) {
    if (m == 0) return;
    k_snp_consts<<<parts_for(m), kBlock, 0, s>>>(freq, m, twof, inv);
    LAUNCH_CHECK();
}

void launch_gdiag(hipStream_t s, const uint8_t* ind, int64_t pitch_i, int32_t n, int64_t m,
                  const double* twof, const double* inv, double* gdiag) {
    k_gdiag<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(ind, pitch_i, n, m, twof, inv,
                                                                     gdiag);
    LAUNCH_CHECK();
}

void launch_rowdot(hipStream_t s, const uint8_t* rows, int64_t pitch, int64_t nrows, int64_t nk,
                   const double* v, int64_t ldv, int k, double* out, int64_t ldo, const int* skip) {
    if (nrows == 0) return;
    const unsigned g = (unsigned)((nrows + kBlock - 1) / kBlock);
    int c = 0;
    while (c < k) {
        const int left = k - c;
        if (left >= 8) {
            k_rowdot<8><<<g, kBlock, 0, s>>>(rows, pitch, nrows, nk, v + c * ldv, ldv, out + c * ldo, ldo, skip);
            c += 8;
        } else if (left >= 4) {
            k_rowdot<4><<<g, kBlock, 0, s>>>(rows, pitch, nrows, nk, v + c * ldv, ldv, out + c * ldo, ldo, skip);
            c += 4;
        } else if (left >= 2) {
            k_rowdot<2><<<g, kBlock, 0, s>>>(rows, pitch, nrows, nk, v + c * ldv, ldv, out + c * ldo, ldo, skip);
            c += 2;
        } else {
            k_rowdot<1><<<g, kBlock, 0, s>>>(rows, pitch, nrows, nk, v + c * ldv, ldv, out + c * ldo, ldo, skip);
            c += 1;
        }
        LAUNCH_CHECK();
    }
}

void launch_colsum_parts(hipStream_t s, int64_t rows, const double* v, int64_t ldv, int k,
                         double* part, const int* skip) {
    dim3 g(parts_for(rows), k);
    k_colsum_parts<<<g, kBlock, 0, s>>>(rows, v, ldv, part, skip);
    LAUNCH_CHECK();
}

void launch_center_snp(hipStream_t s, int64_t m, int k, const double* P1, int64_t ld1,
                       const double* T, const double* twof, const double* inv,
                       const double* theta_cols, double inv_m, double* u, int64_t ldu,
                       double* part_c, const int* skip) {
    dim3 g(parts_for(m), k);
    k_center_snp<<<g, kBlock, 0, s>>>(m, P1, ld1, T, twof, inv, theta_cols, inv_m, u, ldu, part_c,
                                      skip);
    LAUNCH_CHECK();
}

void launch_center_ind(hipStream_t s, int32_t n, int k, const double* P2, int64_t ld2,
                       const double* Csum, double* GZ, int64_t ldg, const int* skip) {
    dim3 g(parts_for(n), k);
    k_center_ind<<<g, kBlock, 0, s>>>(n, P2, ld2, Csum, GZ, ldg, skip);
    LAUNCH_CHECK();
}

}  // namespace globe
