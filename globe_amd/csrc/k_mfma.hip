// Packed-genotype contraction on the int8 matrix cores (v_mfma_i32_16x16x64_i8).
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
//   L = 4: 30-bit fixed point (GLOBE_PREC_MIXED), half the MFMA work of L = 7... of
//          the same cost as BF16 hi+lo pieces, with exact accumulation.
//
// Centering and scaling (w = (g - 2f) inv) are applied afterwards as the FP64
// rank-one correction sum_i w_is z_i = inv_s (P_s - 2 f_s T), T = sum_i z_i computed
// from the quantized z, so the operator is applied exactly to the quantized vector.
//
// Operand layouts (16x16x64 i8, lane l, group g = l / 16):
//   A: row l % 16, 16 K-slots; B: column l % 16, the same 16 K-slots;
//   D: element v of lane l is (row 4 g + v, column l % 16).
// A lane of row fragment rt loads 16 packed bytes (64 genotypes) at byte
// 64 ss + 16 g of its row per super-step ss (4 K-steps); K-step s uses dword s.
// The dword's 16 codes are decoded with two masks and four v_perm_b32 byte-table
// lookups, which yields genotype order {0,4,8,12, 1,5,9,13, 2,6,10,14, 3,7,11,15};
// the digit matrix Bq is stored in exactly that order:
//   Bq[ks][jf][lane][16 bytes], ks = K-step, jf = 16-column fragment.
#include "context.h"

namespace globe {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

// dosage byte table for codes 0..3: 00 -> 2, 01 (missing) -> 0, 10 -> 1, 11 -> 0
constexpr uint32_t kDosTable = 0x00010002u;

__device__ inline v4i decode16(uint32_t w) {
    const uint32_t e = w & 0x33333333u;
    const uint32_t o = (w >> 2) & 0x33333333u;
    v4i d;
    d[0] = (int)__builtin_amdgcn_perm(0u, kDosTable, e & 0x0F0F0F0Fu);
    d[1] = (int)__builtin_amdgcn_perm(0u, kDosTable, o & 0x0F0F0F0Fu);
    d[2] = (int)__builtin_amdgcn_perm(0u, kDosTable, (e >> 4) & 0x0F0F0F0Fu);
    d[3] = (int)__builtin_amdgcn_perm(0u, kDosTable, (o >> 4) & 0x0F0F0F0Fu);
    return d;
}

// RT row fragments x JF column fragments per wave; 4 waves per block.
template <int RT, int JF>
__global__ __launch_bounds__(256) void k_geno_mfma(const uint8_t* __restrict__ rows, int64_t pitch,
                                                   int64_t nrows, int64_t n_ss, int64_t ss_per_split,
                                                   const uint8_t* __restrict__ Bq,
                                                   int32_t* __restrict__ out, int64_t out_rows,
                                                   const int* skip) {
    if (skip && *skip) return;
    __shared__ __attribute__((aligned(16))) uint8_t lds[4 * JF * 1024];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, rr = lane & 15;
    const int64_t row_base = (int64_t)blockIdx.x * (64 * RT) + wave * (16 * RT);
    const int64_t ss0 = (int64_t)blockIdx.y * ss_per_split;
    int64_t ss1 = ss0 + ss_per_split;
    if (ss1 > n_ss) ss1 = n_ss;

    v4i acc[RT][JF];
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
        for (int b = 0; b < JF; ++b) acc[a][b] = v4i{0, 0, 0, 0};

    const uint8_t* rp[RT];
    bool rv[RT];
#pragma unroll
    for (int a = 0; a < RT; ++a) {
        const int64_t r = row_base + a * 16 + rr;
        rv[a] = r < nrows;
        rp[a] = rows + (rv[a] ? r : 0) * pitch + 16 * g;
    }

    uint4 araw[RT];
#pragma unroll
    for (int a = 0; a < RT; ++a)
        araw[a] = (rv[a] && ss0 < ss1) ? *reinterpret_cast<const uint4*>(rp[a] + ss0 * 64)
                                       : make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);

    for (int64_t ss = ss0; ss < ss1; ++ss) {
        __syncthreads();
        {
            const uint4* src = reinterpret_cast<const uint4*>(Bq + ss * (4 * JF * 1024));
            uint4* dst = reinterpret_cast<uint4*>(lds);
#pragma unroll
            for (int q = 0; q < JF; ++q) dst[q * 256 + threadIdx.x] = src[q * 256 + threadIdx.x];
        }
        __syncthreads();
        uint4 cur[RT];
#pragma unroll
        for (int a = 0; a < RT; ++a) cur[a] = araw[a];
        // prefetch the next super-step's packed genotypes
        if (ss + 1 < ss1) {
#pragma unroll
            for (int a = 0; a < RT; ++a)
                if (rv[a]) araw[a] = *reinterpret_cast<const uint4*>(rp[a] + (ss + 1) * 64);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            v4i av[RT];
#pragma unroll
            for (int a = 0; a < RT; ++a) {
                const uint32_t w = s == 0 ? cur[a].x : s == 1 ? cur[a].y : s == 2 ? cur[a].z : cur[a].w;
                av[a] = decode16(w);
            }
#pragma unroll
            for (int b = 0; b < JF; ++b) {
                const v4i bv = *reinterpret_cast<const v4i*>(lds + ((s * JF + b) * 64 + lane) * 16);
#pragma unroll
                for (int a = 0; a < RT; ++a)
                    acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[a], bv, acc[a][b], 0, 0, 0);
            }
        }
    }
    // D element v of lane -> (row 4 g + v, column rr)
    int32_t* o = out + (int64_t)blockIdx.y * out_rows * (JF * 16);
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t r = row_base + a * 16 + 4 * g + v;
            if (r < nrows)
#pragma unroll
                for (int b = 0; b < JF; ++b) o[r * (JF * 16) + b * 16 + rr] = acc[a][b][v];
        }
}

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

// scale exponent per column: sh_c = (8L - 2) - e_c with max|v| < 2^e_c
__global__ void k_scales(const double* __restrict__ part, int nparts, int k, int limbs, int* sh,
                         const int* skip) {
    if (skip && *skip) return;
    const int c = threadIdx.x;
    if (c >= k) return;
    double m = 0.0;
    for (int j = 0; j < nparts; ++j) m = fmax(m, part[c * kMaxParts + j]);
    int e = 0;
    if (m > 0.0) (void)frexp(m, &e);  // m < 2^e
    sh[c] = (m > 0.0 && isfinite(m)) ? (8 * limbs - 2) - e : 0;
}

// Quantize v (k columns, K rows valid, Kpad rows total) into the digit layout and
// emit per-block partials of sum_kappa w_kappa * q_kappa 2^-sh (T for pass 1 with
// w = 1, C for pass 2 with w = 2 f).
__global__ __launch_bounds__(256) void k_quantize(const double* __restrict__ v, int64_t ldv,
                                                  int64_t K, int64_t Kpad, int k, int limbs, int JF,
                                                  const int* __restrict__ sh,
                                                  const double* __restrict__ wts,
                                                  uint8_t* __restrict__ Bq, double* part,
                                                  const int* skip) {
    if (skip && *skip) return;
    const int j16 = blockIdx.y;
    const int c = j16 / limbs, l = j16 % limbs;
    const bool live = c < k;
    double acc = 0.0;
    for (int64_t kap = blockIdx.x * 256ll + threadIdx.x; kap < Kpad; kap += (int64_t)gridDim.x * 256) {
        int dig = 0;
        if (live && kap < K) {
            const double z = v[c * ldv + kap];
            long long q = llrint(ldexp(z, sh[c]));
            if (l == 0 && wts) acc += wts[kap] * ldexp((double)q, -sh[c]);
            if (l == 0 && !wts) acc += ldexp((double)q, -sh[c]);
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
        const int gg = rem >> 6, s = (rem >> 4) & 3, w = rem & 15;
        const int qslot = 4 * (w & 3) + (w >> 2);
        const int64_t ks = 4 * ss + s;
        const int jf = j16 >> 4, jj = j16 & 15;
        Bq[((ks * JF + jf) * 64 + (16 * gg + jj)) * 16 + qslot] = (uint8_t)(int8_t)dig;
    }
    if (live && l == 0) {
        const double r = block_sum(acc);
        if (threadIdx.x == 0) part[c * kMaxParts + blockIdx.x] = r;
    }
}

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

// pass-1 epilogue: P = combine * 2^-sh; u = sig2g / M * inv^2 * (P - 2 f T); colmax of u
__global__ __launch_bounds__(256) void k_epi1(const int32_t* __restrict__ out, int nsplit,
                                              int64_t m, int J16, int limbs,
                                              const int* __restrict__ sh,
                                              const double* __restrict__ T,
                                              const double* __restrict__ twof,
                                              const double* __restrict__ inv,
                                              const double* __restrict__ th, double inv_m,
                                              double* __restrict__ u, int64_t ldu, double* part,
                                              const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    const double scale = th[c * 5 + 1] * inv_m, tc = T[c];
    const int shc = sh[c];
    double mx = 0.0;
    for (int64_t s = blockIdx.x * 256ll + threadIdx.x; s < m; s += (int64_t)gridDim.x * 256) {
        const double P = ldexp(combine(out, nsplit, m * J16, s * J16 + c * limbs, limbs), -shc);
        const double iv = inv[s];
        const double uv = scale * iv * iv * (P - twof[s] * tc);
        u[c * ldu + s] = uv;
        mx = fmax(mx, fabs(uv));
    }
    const double r = block_max(mx);
    if (threadIdx.x == 0) part[c * kMaxParts + blockIdx.x] = r;
}

// pass-2 epilogue: GZ = combine * 2^-sh - C
__global__ __launch_bounds__(256) void k_epi2(const int32_t* __restrict__ out, int nsplit,
                                              int64_t n, int J16, int limbs,
                                              const int* __restrict__ sh,
                                              const double* __restrict__ Cs,
                                              double* __restrict__ GZ, int64_t ldg,
                                              const int* skip) {
    if (skip && *skip) return;
    const int c = blockIdx.y;
    const int shc = sh[c];
    const double cc = Cs[c];
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        GZ[c * ldg + i] = ldexp(combine(out, nsplit, n * J16, i * J16 + c * limbs, limbs), -shc) - cc;
}

__global__ void k_reduce_sum(const double* __restrict__ part, int nparts, int k, double* out,
                             const int* skip) {
    if (skip && *skip) return;
    const int c = threadIdx.x;
    if (c >= k) return;
    double s = 0.0;
    for (int j = 0; j < nparts; ++j) s += part[c * kMaxParts + j];
    out[c] = s;
}

struct Launch {
    int rt, jf;
};

int jf_for(int J) {
    const int f = (J + 15) / 16;
    return f <= 1 ? 1 : f <= 2 ? 2 : f <= 4 ? 4 : 8;
}

void launch_mfma(hipStream_t s, int RT, int JF, const uint8_t* rows, int64_t pitch, int64_t nrows,
                 int64_t n_ss, int nsplit, const uint8_t* Bq, int32_t* out, const int* skip) {
    const int64_t per = (n_ss + nsplit - 1) / nsplit;
    dim3 grid((unsigned)((nrows + 64 * RT - 1) / (64 * RT)), (unsigned)nsplit);
#define GLOBE_MFMA_CASE(R, F)                                                                 \
    if (RT == R && JF == F) {                                                                 \
        k_geno_mfma<R, F><<<grid, 256, 0, s>>>(rows, pitch, nrows, n_ss, per, Bq, out, nrows, \
                                               skip);                                          \
        HIP_CHECK(hipGetLastError());                                                         \
        return;                                                                               \
    }
    GLOBE_MFMA_CASE(4, 1)
    GLOBE_MFMA_CASE(4, 2)
    GLOBE_MFMA_CASE(2, 4)
    GLOBE_MFMA_CASE(2, 8)
    GLOBE_MFMA_CASE(2, 1)
    GLOBE_MFMA_CASE(2, 2)
    GLOBE_MFMA_CASE(1, 4)
    GLOBE_MFMA_CASE(1, 8)
#undef GLOBE_MFMA_CASE
    throw Error(GLOBE_E_INVALID, "no MFMA tile for this shape");
}

int rt_for(int JF) { return JF <= 2 ? 4 : 2; }

// split K so that the grid has at least ~2048 blocks, keeping >= 4 super-steps each
int split_for(int64_t nrows, int RT, int64_t n_ss) {
    const int64_t rb = (nrows + 64 * RT - 1) / (64 * RT);
    int64_t sp = (2048 + rb - 1) / rb;
    const int64_t maxsp = n_ss / 4 > 1 ? n_ss / 4 : 1;
    if (sp > maxsp) sp = maxsp;
    if (sp > 64) sp = 64;
    if (sp < 1) sp = 1;
    return (int)sp;
}

}  // namespace

// GZ (n x k) = sig2g G S0 for the packed-genotype cohort with `limbs` int8 digits.
void mfma_geno_product(globe_ctx* c, const double* S0, int64_t ldn, int k, double* GZ, int64_t ldg,
                       const int* skip, int limbs, const double* theta_cols) {
    hipStream_t s = c->stream;
    const int64_t n = c->n, m = c->m;
    const int J = k * limbs;
    const int JF = jf_for(J);
    require(J <= 128, GLOBE_E_INVALID, "k * limbs must be <= 128 for the MFMA contraction");
    const int RT = rt_for(JF);
    const int64_t Kpad1 = c->pitch_s * 4, Kpad2 = c->pitch_i * 4;  // genotypes per padded row
    const int64_t nss1 = Kpad1 / 256, nss2 = Kpad2 / 256;
    const int sp1 = split_for(m, RT, nss1), sp2 = split_for(n, RT, nss2);
    const size_t bq1 = (size_t)Kpad1 * JF * 16, bq2 = (size_t)Kpad2 * JF * 16;
    const size_t o1 = (size_t)sp1 * m * JF * 16 * 4, o2 = (size_t)sp2 * n * JF * 16 * 4;
    // scratch: Bq (max of both passes) | out32 (max) | sh (2 x kMaxCols ints)
    const size_t bq = bq1 > bq2 ? bq1 : bq2, ob = o1 > o2 ? o1 : o2;
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

    // pass 1: P1[s][c] = sum_i g_is S0[c][i] on the SNP-major rows
    const int npn = parts_for(n);
    k_colmax_parts<<<dim3(npn, k), 256, 0, s>>>(n, S0, ldn, part, skip);
    k_scales<<<1, kMaxCols, 0, s>>>(part, npn, k, limbs, sh1, skip);
    const int qg1 = parts_for(Kpad1);
    k_quantize<<<dim3(qg1, JF * 16), 256, 0, s>>>(S0, ldn, n, Kpad1, k, limbs, JF, sh1, nullptr, Bq,
                                                  part, skip);
    k_reduce_sum<<<1, kMaxCols, 0, s>>>(part, qg1, k, c->w_T.as<double>(), skip);
    HIP_CHECK(hipGetLastError());
    // algorithmic work per pass: rows x ceil(K/4) packed bytes, 2 N M k flops
    const double flops = 2.0 * (double)n * (double)m * k;
    auto timed = [&](int pass, int64_t rows, int64_t K, auto&& fn) {
        if (!c->prof_on) {
            fn();
            return;
        }
        globe_ctx::ProfEvent e{};
        HIP_CHECK(hipEventCreate(&e.a));
        HIP_CHECK(hipEventCreate(&e.b));
        e.pass = pass;
        e.iter = c->prof_iter;
        e.bytes = (double)rows * (double)((K + 3) / 4);
        e.flops = flops;
        e.issued = 2.0 * (double)rows * (double)(pass == 0 ? Kpad1 : Kpad2) * (JF * 16);
        HIP_CHECK(hipEventRecord(e.a, s));
        fn();
        HIP_CHECK(hipEventRecord(e.b, s));
        c->prof.push_back(e);
    };
    timed(0, m, n, [&] { launch_mfma(s, RT, JF, c->d_snp.as<uint8_t>(), c->pitch_s, m, nss1, sp1, Bq, out, skip); });
    const int npm = parts_for(m);
    k_epi1<<<dim3(npm, k), 256, 0, s>>>(out, sp1, m, JF * 16, limbs, sh1, c->w_T.as<double>(),
                                         c->d_twof.as<double>(), c->d_inv.as<double>(), th,
                                         1.0 / (double)m, U, m, part, skip);
    HIP_CHECK(hipGetLastError());
    // pass 2: GZ[c][i] = sum_s g_is u[c][s] - C[c] on the individual-major rows
    k_scales<<<1, kMaxCols, 0, s>>>(part, npm, k, limbs, sh2, skip);
    const int qg2 = parts_for(Kpad2);
    k_quantize<<<dim3(qg2, JF * 16), 256, 0, s>>>(U, m, m, Kpad2, k, limbs, JF, sh2,
                                                  c->d_twof.as<double>(), Bq, part, skip);
    k_reduce_sum<<<1, kMaxCols, 0, s>>>(part, qg2, k, c->w_C.as<double>(), skip);
    HIP_CHECK(hipGetLastError());
    timed(1, n, m, [&] { launch_mfma(s, RT, JF, c->d_ind.as<uint8_t>(), c->pitch_i, n, nss2, sp2, Bq, out, skip); });
    k_epi2<<<dim3(npn, k), 256, 0, s>>>(out, sp2, n, JF * 16, limbs, sh2, c->w_C.as<double>(), GZ,
                                         ldg, skip);
    HIP_CHECK(hipGetLastError());
}

}  // namespace globe
