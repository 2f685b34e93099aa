/*
 * globe_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * FP64 CPU restatement of the reference's LGMM solve. Each function cites the
 * reference lines it restates (paths relative to the reference repository root).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library; libglobe_amd.so never links it.
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   orc_decode / orc_bed_counts / orc_bed_qc : PLINK 1 .bed spec brute force in
 *       numpy, phantom (n % 4 != 0) cases, and the compiled reference runPCGsolver.
 *   orc_vx / orc_vx_dense / orc_genetic_diag   : an explicit dense V built in numpy
 *       from W W^T / M, symmetry, and the compiled reference's outputs.
 *   orc_precond_*                              : numpy dense inverse, Q Q^-1 = I.
 *   orc_pcg / orc_gls_blup / orc_bhat          : numpy dense solve + GLS/Henderson
 *       closed forms, and golden _beta/_uhat/_bhat.bin files written by the
 *       compiled reference runPCGsolver (FP32, tol 1e-5, so ~1e-4 agreement).
 *   orc_f_theta / orc_aireml_eval              : exact dense P, f = dV P y,
 *       tr(P dV) with the same probes, AI = f^T P f / 2 (numpy).
 *   orc_tildeV_solve / orc_sample_tildeV       : numpy dense block solve / the
 *       covariance of the sampling formula.
 *   orc_vx_ref32                               : agrees with orc_vx to FP32 rounding.
 */
#include "globe_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* PCGsolver.cpp:83-94 (decode2Bits) -- 00 -> 2 (A1 homozygote), 10 -> 1,
 * 11 -> 0, 01 -> missing (-1). The 2-bit code of individual i in a SNP row is
 * (byte[i/4] >> 2*(i%4)) & 3 (PCGsolver.cpp:825-827). */
int orc_decode(uint8_t code) {
    switch (code & 3) {
        case 0: return 2;
        case 2: return 1;
        case 3: return 0;
        default: return -1;
    }
}

static inline int geno_at(const uint8_t* row, int64_t i) {
    return orc_decode((uint8_t)(row[i >> 2] >> (2 * (i & 3))));
}

/* PCGsolver.cpp:706-724: per-SNP A1 allele count (2*#00 + #10) and missing
 * count (#01) over the first n individuals only (phantom pairs of the last byte
 * are excluded, as subtractPhantomGenotypes does at :590-613). */
void orc_bed_counts(const uint8_t* rows, int64_t m, int32_t n, int64_t pitch,
                    int64_t* alt, int64_t* miss) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < m; ++s) {
        const uint8_t* row = rows + s * pitch;
        int64_t a = 0, mi = 0;
        for (int64_t i = 0; i < n; ++i) {
            int g = geno_at(row, i);
            if (g < 0) mi++; else a += g;
        }
        alt[s] = a;
        miss[s] = mi;
    }
}

/* PCGsolver.cpp:726-743: the keep decision in FP32, bit for bit:
 *   n_obs = N - miss; maf = (float)alt / (2.0f * n_obs)  (0 if n_obs == 0)
 *   missRate = (float)miss / (float)N
 *   keep iff maf >= minMAF && maf <= 1 - minMAF && missRate <= maxMiss
 * Returns the number of kept SNPs; maf_out/miss_rate_out are per input SNP. */
int64_t orc_bed_qc(const uint8_t* rows, int64_t m, int32_t n, int64_t pitch,
                   float min_maf, float max_miss, uint8_t* keep, float* maf_out,
                   float* miss_rate_out) {
    int64_t* alt = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    int64_t* miss = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    orc_bed_counts(rows, m, n, pitch, alt, miss);
    int64_t kept = 0;
    for (int64_t s = 0; s < m; ++s) {
        int n_obs = n - (int)miss[s];
        volatile float denom = 2.0f * (float)n_obs;
        float maf = (n_obs > 0) ? (float)alt[s] / denom : 0.0f;
        float mr = (float)miss[s] / (float)n;
        int k = (maf >= min_maf && maf <= (1.0f - min_maf) && mr <= max_miss);
        keep[s] = (uint8_t)k;
        if (maf_out) maf_out[s] = maf;
        if (miss_rate_out) miss_rate_out[s] = mr;
        kept += k;
    }
    free(alt);
    free(miss);
    return kept;
}

/* PCGsolver.cpp:838-841 restated in FP64 from the stored FP32 frequency:
 * inv_std = 1/sqrt(2 f (1-f)), or 1 when 2 f (1-f) <= 1e-12. */
double orc_inv_std(float f) {
    double fd = (double)f;
    double d = 2.0 * fd * (1.0 - fd);
    return (d > 1e-12) ? 1.0 / sqrt(d) : 1.0;
}

/* standardized genotype w = (g - 2f) inv_std, missing g := 0 (PCGsolver.cpp:831-842) */
static inline double std_geno(const orc_model* md, int64_t s, int64_t i, const double* inv) {
    int g = geno_at(md->geno + s * md->pitch, i);
    if (g < 0) g = 0;
    return ((double)g - 2.0 * (double)md->freq[s]) * inv[s];
}

static double* inv_std_table(const orc_model* md) {
    double* inv = (double*)malloc(sizeof(double) * (size_t)(md->m > 0 ? md->m : 1));
    for (int64_t s = 0; s < md->m; ++s) inv[s] = orc_inv_std(md->freq[s]);
    return inv;
}

static inline double grm_at(const float* lt, int64_t i, int64_t j) {
    if (j > i) { int64_t t = i; i = j; j = t; }
    return (double)lt[i * (i + 1) / 2 + j];
}

/* PCGsolver.cpp:925-989 (computeGeneticDiagonal): diag(G) = (1/M) sum_s w_is^2,
 * or the GRM diagonal (:936-945). */
void orc_genetic_diag(const orc_model* md, double* gdiag) {
    const int64_t n = md->n;
    if (md->use_grm) {
        for (int64_t i = 0; i < n; ++i) gdiag[i] = grm_at(md->grm_lt, i, i);
        return;
    }
    if (md->m == 0) { memset(gdiag, 0, sizeof(double) * (size_t)n); return; }
    double* inv = inv_std_table(md);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int64_t s = 0; s < md->m; ++s) {
            double w = std_geno(md, s, i, inv);
            acc += w * w;
        }
        gdiag[i] = acc / (double)md->m;
    }
    free(inv);
}

/* GZ = sig2g * G * Z for Z [k][n]; geno: G = W W^T / M (PCGsolver.cpp:1130-1201),
 * GRM: lower-triangle lookup on both triangles (:1202-1213). */
static void genetic_product(const orc_model* md, const double* Z, int k, double* GZ) {
    const int64_t n = md->n;
    const double sg = md->theta[1];
    memset(GZ, 0, sizeof(double) * (size_t)n * k);
    if (md->use_grm) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            for (int c = 0; c < k; ++c) {
                double acc = 0.0;
                for (int64_t j = 0; j < n; ++j) acc += grm_at(md->grm_lt, i, j) * Z[c * n + j];
                GZ[c * n + i] = sg * acc;
            }
        }
        return;
    }
    const int64_t m = md->m;
    if (m == 0) return;
    double* inv = inv_std_table(md);
    double* part = (double*)calloc((size_t)m * k, sizeof(double));
    /* pass 1: part[s] = sum_i w_is Z_i (:1143-1171) */
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < m; ++s) {
        for (int c = 0; c < k; ++c) {
            double acc = 0.0;
            for (int64_t i = 0; i < n; ++i) acc += std_geno(md, s, i, inv) * Z[c * n + i];
            part[c * m + s] = acc;
        }
    }
    /* pass 2: GZ_i = (sig2g / M) sum_s w_is part[s] (:1173-1199) */
    const double scale = sg / (double)m;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        for (int c = 0; c < k; ++c) {
            double acc = 0.0;
            for (int64_t s = 0; s < m; ++s) acc += std_geno(md, s, i, inv) * part[c * m + s];
            GZ[c * n + i] = scale * acc;
        }
    }
    free(part);
    free(inv);
}

/* PCGsolver.cpp:1004-1011: block starts where indiv changes; returns #blocks. */
int64_t orc_blocks(const orc_model* md, int64_t* starts) {
    int64_t nb = 0;
    starts[nb++] = 0;
    for (int64_t r = 1; r < md->nrec; ++r)
        if (md->indiv[r] != md->indiv[r - 1]) starts[nb++] = r;
    starts[nb] = md->nrec;
    return nb;
}

/* PCGsolver.cpp:1095-1266 (computeVx_multi), FP64:
 *   Z = J^T x; GZ = sig2g G Z; (grm_only: return GZ, the multGRM_uhat mode)
 *   y = J GZ + blockwise (sig2b0 + sigb01 (t_i + t_j) + sig2b1 t_i t_j) x_j + sig2e x */
void orc_vx(const orc_model* md, const double* x, int k, double* y, int grm_only) {
    const int64_t n = md->n, nrec = md->nrec;
    double* Z = (double*)calloc((size_t)n * k, sizeof(double));
    double* GZ = (double*)malloc(sizeof(double) * (size_t)n * k);
    for (int c = 0; c < k; ++c)
        for (int64_t r = 0; r < nrec; ++r) Z[c * n + md->indiv[r]] += x[c * nrec + r];
    genetic_product(md, Z, k, GZ);
    if (grm_only) {
        memcpy(y, GZ, sizeof(double) * (size_t)n * k);
        free(Z); free(GZ);
        return;
    }
    for (int c = 0; c < k; ++c)
        for (int64_t r = 0; r < nrec; ++r) y[c * nrec + r] = GZ[c * n + md->indiv[r]];
    const double sb0 = md->theta[2], sb01 = md->theta[3], sb1 = md->theta[4], se = md->theta[0];
    int64_t* st = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nrec + 1));
    int64_t nb = orc_blocks(md, st);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; ++b) {
        for (int64_t i = st[b]; i < st[b + 1]; ++i) {
            double ti = md->times[i];
            for (int64_t j = st[b]; j < st[b + 1]; ++j) {
                double tj = md->times[j];
                double v = sb0 + sb01 * (ti + tj) + sb1 * (ti * tj);
                for (int c = 0; c < k; ++c) y[c * nrec + i] += v * x[c * nrec + j];
            }
        }
    }
    for (int c = 0; c < k; ++c)
        for (int64_t r = 0; r < nrec; ++r) y[c * nrec + r] += se * x[c * nrec + r];
    free(st); free(Z); free(GZ);
}

/* varCompEst.cpp:309-352 (computeV_ij / computeVx_multi_VC): the same V formed
 * entry by entry over all record pairs; G from the GRM or from W W^T / M. */
void orc_vx_dense(const orc_model* md, const double* x, int k, double* y) {
    const int64_t n = md->n, nrec = md->nrec;
    double* G = (double*)malloc(sizeof(double) * (size_t)n * n);
    if (md->use_grm) {
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = 0; j < n; ++j) G[i * n + j] = grm_at(md->grm_lt, i, j);
    } else {
        double* inv = inv_std_table(md);
        double* W = (double*)malloc(sizeof(double) * (size_t)n * (md->m > 0 ? md->m : 1));
        for (int64_t i = 0; i < n; ++i)
            for (int64_t s = 0; s < md->m; ++s) W[i * md->m + s] = std_geno(md, s, i, inv);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = 0; j < n; ++j) {
                double acc = 0.0;
                for (int64_t s = 0; s < md->m; ++s) acc += W[i * md->m + s] * W[j * md->m + s];
                G[i * n + j] = md->m > 0 ? acc / (double)md->m : 0.0;
            }
        free(W); free(inv);
    }
    const double se = md->theta[0], sg = md->theta[1], sb0 = md->theta[2], sb01 = md->theta[3],
                 sb1 = md->theta[4];
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < nrec; ++r) {
        int32_t ir = md->indiv[r];
        double tr = md->times[r];
        for (int c = 0; c < k; ++c) y[c * nrec + r] = 0.0;
        for (int64_t s = 0; s < nrec; ++s) {
            int32_t is = md->indiv[s];
            double ts = md->times[s];
            double v = sg * G[(int64_t)ir * n + is];
            if (ir == is) v += sb0 + sb01 * (tr + ts) + sb1 * tr * ts;
            if (r == s) v += se;
            for (int c = 0; c < k; ++c) y[c * nrec + r] += v * x[c * nrec + s];
        }
    }
    free(G);
}

/* out_i = sum_s w_is a_s (standardized genotypes times a SNP-space vector;
 * used by the synthetic phenotype recipe u = sqrt(sig2g/M) W a). */
void orc_geno_Wa(const orc_model* md, const double* a, double* out) {
    double* inv = inv_std_table(md);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < md->n; ++i) {
        double acc = 0.0;
        for (int64_t s = 0; s < md->m; ++s) acc += std_geno(md, s, i, inv) * a[s];
        out[i] = acc;
    }
    free(inv);
}

/* FP32 OpenMP restatement of computeVx_multi (PCGsolver.cpp:1095-1266) kept as the
 * CPU timing baseline: same per-element getGenotype addressing (chunk = s / perChunk,
 * offset = s % perChunk, :817-827), per-element sqrt/divide (:1159-1162), 50,000-SNP
 * chunks (:1135), thread-local accumulators, explicit n_b^2 time loop (:1237-1254). */
void orc_vx_ref32(const orc_model* md, const float* x, int k, float* y, int grm_only) {
    const int N = md->n;
    const int64_t nrec = md->nrec;
    const int64_t out_rows = grm_only ? N : nrec;
    for (int64_t q = 0; q < out_rows * k; ++q) y[q] = 0.0f;
    float* Z = (float*)calloc((size_t)N * k, sizeof(float));
    float* GZ = (float*)calloc((size_t)N * k, sizeof(float));
    for (int64_t r = 0; r < nrec; ++r)
        for (int c = 0; c < k; ++c) Z[(int64_t)c * N + md->indiv[r]] += x[c * nrec + r];
    const float sg = (float)md->theta[1];
    if (!md->use_grm) {
        const size_t num_snps = (size_t)md->m;
        const float inv_num_snp = num_snps > 0 ? 1.0f / (float)num_snps : 0.0f;
        size_t per_chunk = (size_t)(1e9 / (double)md->pitch);
        if (per_chunk == 0) per_chunk = 50000;
        size_t n_chunks = (num_snps + per_chunk - 1) / per_chunk;
        const uint8_t** chunks = (const uint8_t**)malloc(sizeof(void*) * (n_chunks + 1));
        for (size_t c = 0; c < n_chunks; ++c) chunks[c] = md->geno + c * per_chunk * md->pitch;
        const size_t CH = 50000;
        for (size_t s0 = 0; s0 < num_snps; s0 += CH) {
            size_t s1 = s0 + CH < num_snps ? s0 + CH : num_snps;
            float* part = (float*)calloc((s1 - s0) * (size_t)k, sizeof(float));
#pragma omp parallel
            {
                float* acc = (float*)malloc(sizeof(float) * (size_t)k);
#pragma omp for schedule(dynamic)
                for (long s = (long)s0; s < (long)s1; ++s) {
                    for (int c = 0; c < k; ++c) acc[c] = 0.0f;
                    for (int i = 0; i < N; ++i) {
                        int ci = (int)((size_t)s / per_chunk);
                        int off = (int)((size_t)s % per_chunk);
                        uint8_t code = (chunks[ci][(size_t)off * md->pitch + i / 4] >> ((i % 4) * 2)) & 3;
                        int g = orc_decode(code);
                        if (g < 0) g = 0;
                        float maf = md->freq[s];
                        float denom = 2.0f * maf * (1.0f - maf);
                        float inv = (denom > 1e-12f) ? (1.0f / sqrtf(denom)) : 1.0f;
                        float w = ((float)g - 2.0f * maf) * inv;
                        for (int c = 0; c < k; ++c) acc[c] += w * Z[(int64_t)c * N + i];
                    }
                    for (int c = 0; c < k; ++c) part[(size_t)(s - (long)s0) * k + c] = acc[c];
                }
                free(acc);
            }
#pragma omp parallel
            {
                float* res = (float*)malloc(sizeof(float) * (size_t)k);
#pragma omp for schedule(dynamic)
                for (int i = 0; i < N; ++i) {
                    for (int c = 0; c < k; ++c) res[c] = 0.0f;
                    for (size_t s = s0; s < s1; ++s) {
                        int ci = (int)(s / per_chunk);
                        int off = (int)(s % per_chunk);
                        uint8_t code = (chunks[ci][(size_t)off * md->pitch + i / 4] >> ((i % 4) * 2)) & 3;
                        int g = orc_decode(code);
                        if (g < 0) g = 0;
                        float maf = md->freq[s];
                        float denom = 2.0f * maf * (1.0f - maf);
                        float inv = (denom > 1e-12f) ? (1.0f / sqrtf(denom)) : 1.0f;
                        float w = ((float)g - 2.0f * maf) * inv;
                        for (int c = 0; c < k; ++c) res[c] += w * part[(s - s0) * (size_t)k + c];
                    }
                    for (int c = 0; c < k; ++c) GZ[(int64_t)c * N + i] += sg * inv_num_snp * res[c];
                }
                free(res);
            }
            free(part);
        }
        free(chunks);
    } else {
#pragma omp parallel for schedule(dynamic)
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                float gij = (float)grm_at(md->grm_lt, i, j);
                for (int c = 0; c < k; ++c) GZ[(int64_t)c * N + i] += sg * gij * Z[(int64_t)c * N + j];
            }
    }
    if (grm_only) {
        memcpy(y, GZ, sizeof(float) * (size_t)N * k);
        free(Z); free(GZ);
        return;
    }
    for (int64_t r = 0; r < nrec; ++r)
        for (int c = 0; c < k; ++c) y[c * nrec + r] += GZ[(int64_t)c * N + md->indiv[r]];
    int64_t* st = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nrec + 1));
    int64_t nb = orc_blocks(md, st);
    const float sb0 = (float)md->theta[2], sb01 = (float)md->theta[3], sb1 = (float)md->theta[4];
#pragma omp parallel for
    for (long b = 0; b < (long)nb; ++b)
        for (int64_t i = st[b]; i < st[b + 1]; ++i) {
            float ti = (float)md->times[i];
            for (int64_t j = st[b]; j < st[b + 1]; ++j) {
                float tj = (float)md->times[j];
                float v = sb0 + sb01 * (ti + tj) + sb1 * (ti * tj);
                for (int c = 0; c < k; ++c) y[c * nrec + i] += v * x[c * nrec + j];
            }
        }
    const float se = (float)md->theta[0];
#pragma omp parallel for
    for (long r = 0; r < (long)nrec; ++r)
        for (int c = 0; c < k; ++c) y[c * nrec + r] += se * x[c * nrec + r];
    free(st); free(Z); free(GZ);
}

/* FP32 OpenMP restatement of computeVx_multi_VC (varCompEst.cpp:322-352) with
 * computeV_ij (:309-319), kept as the CPU timing baseline of runAIREML's operator: for
 * every record pair j <= i, V_ij = sig2g A(indiv_i, indiv_j) from the packed lower
 * triangle (getElement_vecLT, :304-306) + the time term when the records share an
 * individual + sig2e on the diagonal, accumulated into both rows in thread-local
 * result matrices (schedule dynamic over i) that are added at the end. Needs the GRM. */
void orc_vx_vc_ref32(const orc_model* md, const float* x, int k, float* y) {
    const int64_t nrec = md->nrec;
    for (int64_t q = 0; q < nrec * k; ++q) y[q] = 0.0f;
    if (!md->use_grm) return;
    const float sg = (float)md->theta[1], sb0 = (float)md->theta[2], sb01 = (float)md->theta[3],
                sb1 = (float)md->theta[4], se = (float)md->theta[0];
#pragma omp parallel
    {
        float* loc = (float*)calloc((size_t)nrec * k, sizeof(float));
#pragma omp for schedule(dynamic)
        for (long i = 0; i < (long)nrec; ++i) {
            const int ii = md->indiv[i];
            const float ti = (float)md->times[i];
            for (long j = 0; j <= i; ++j) {
                const int ij = md->indiv[j];
                const float tj = (float)md->times[j];
                const size_t a = ii < ij ? (size_t)ij * (ij + 1) / 2 + ii : (size_t)ii * (ii + 1) / 2 + ij;
                float v = sg * md->grm_lt[a];
                v += ii == ij ? sb0 + sb01 * (ti + tj) + sb1 * ti * tj : 0.0f;
                v += i == j ? se : 0.0f;
                for (int c = 0; c < k; ++c) {
                    loc[(size_t)c * nrec + i] += v * x[(size_t)c * nrec + j];
                    if (i != j) loc[(size_t)c * nrec + j] += v * x[(size_t)c * nrec + i];
                }
            }
        }
#pragma omp critical
        for (int64_t q = 0; q < nrec * k; ++q) y[q] += loc[q];
        free(loc);
    }
}

/* Gauss-Jordan inverse with partial pivoting (stands in for Eigen's
 * partialPivLu-based inverse(), thirdParty/eigen/Eigen/src/LU/InverseImpl.h:23-28).
 * Returns 0 on success, 1 if a pivot is exactly zero. */
int orc_inverse(int n, const double* a, double* ainv) {
    double* w = (double*)malloc(sizeof(double) * (size_t)n * 2 * n);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) w[i * 2 * n + j] = a[i * n + j];
        for (int j = 0; j < n; ++j) w[i * 2 * n + n + j] = (i == j) ? 1.0 : 0.0;
    }
    int bad = 0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(w[r * 2 * n + c]) > fabs(w[piv * 2 * n + c])) piv = r;
        if (w[piv * 2 * n + c] == 0.0) { bad = 1; continue; }
        if (piv != c)
            for (int j = 0; j < 2 * n; ++j) {
                double t = w[c * 2 * n + j]; w[c * 2 * n + j] = w[piv * 2 * n + j]; w[piv * 2 * n + j] = t;
            }
        double d = w[c * 2 * n + c];
        for (int j = 0; j < 2 * n; ++j) w[c * 2 * n + j] /= d;
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = w[r * 2 * n + c];
            if (f == 0.0) continue;
            for (int j = 0; j < 2 * n; ++j) w[r * 2 * n + j] -= f * w[c * 2 * n + j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) ainv[i * n + j] = w[i * 2 * n + n + j];
    free(w);
    return bad;
}

int64_t orc_precond_size(const orc_model* md) {
    int64_t* st = (int64_t*)malloc(sizeof(int64_t) * (size_t)(md->nrec + 1));
    int64_t nb = orc_blocks(md, st), tot = 0;
    for (int64_t b = 0; b < nb; ++b) {
        int64_t s = st[b + 1] - st[b];
        tot += s * (s + 1) / 2;
    }
    free(st);
    return tot;
}

/* PCGsolver.cpp:991-1058 (computePreconditioner_Block) / varCompEst.cpp:355-415:
 * Q_b = sig2g G_uu 1 1^T + [sig2b0 + sigb01 (t_i + t_j) + sig2b1 t_i t_j] + sig2e I,
 * explicitly inverted and stored packed lower-triangular, blocks back to back.
 * no_grm = 1 uses G_uu := 1 (tilde-V, varCompEst.cpp:839-858). */
void orc_precond_build(const orc_model* md, const double* gdiag, double* qinv, int no_grm) {
    int64_t* st = (int64_t*)malloc(sizeof(int64_t) * (size_t)(md->nrec + 1));
    int64_t nb = orc_blocks(md, st);
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb + 1));
    off[0] = 0;
    for (int64_t b = 0; b < nb; ++b) {
        int64_t s = st[b + 1] - st[b];
        off[b + 1] = off[b] + s * (s + 1) / 2;
    }
    const double se = md->theta[0], sg = md->theta[1], sb0 = md->theta[2], sb01 = md->theta[3],
                 sb1 = md->theta[4];
#pragma omp parallel for schedule(dynamic)
    for (int64_t b = 0; b < nb; ++b) {
        int64_t s0 = st[b];
        int bs = (int)(st[b + 1] - s0);
        double gval = sg * (no_grm ? 1.0 : gdiag[md->indiv[s0]]);
        double* M = (double*)malloc(sizeof(double) * (size_t)bs * bs);
        double* Mi = (double*)malloc(sizeof(double) * (size_t)bs * bs);
        for (int i = 0; i < bs; ++i)
            for (int j = 0; j <= i; ++j) {
                double ti = md->times[s0 + i], tj = md->times[s0 + j];
                double v = gval + sb0 + sb01 * (ti + tj) + sb1 * (ti * tj);
                if (i == j) v += se;
                M[i * bs + j] = v;
                M[j * bs + i] = v;
            }
        orc_inverse(bs, M, Mi);
        double* dst = qinv + off[b];
        for (int i = 0; i < bs; ++i)
            for (int j = 0; j <= i; ++j) dst[i * (i + 1) / 2 + j] = Mi[i * bs + j];
        free(M); free(Mi);
    }
    free(off); free(st);
}

/* PCGsolver.cpp:1060-1093 (applyPreconditioner): packed symmetric block matvec. */
void orc_precond_apply(const orc_model* md, const double* qinv, const double* r, int k, double* z) {
    const int64_t nrec = md->nrec;
    int64_t* st = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nrec + 1));
    int64_t nb = orc_blocks(md, st);
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb + 1));
    off[0] = 0;
    for (int64_t b = 0; b < nb; ++b) {
        int64_t s = st[b + 1] - st[b];
        off[b + 1] = off[b] + s * (s + 1) / 2;
    }
    for (int c = 0; c < k; ++c) {
        const double* rc = r + (int64_t)c * nrec;
        double* zc = z + (int64_t)c * nrec;
#pragma omp parallel for schedule(static)
        for (int64_t b = 0; b < nb; ++b) {
            int64_t s0 = st[b];
            int bs = (int)(st[b + 1] - s0);
            const double* L = qinv + off[b];
            for (int i = 0; i < bs; ++i) zc[s0 + i] = 0.0;
            for (int i = 0; i < bs; ++i)
                for (int j = 0; j <= i; ++j) {
                    double v = L[i * (i + 1) / 2 + j];
                    zc[s0 + i] += v * rc[s0 + j];
                    if (i != j) zc[s0 + j] += v * rc[s0 + i];
                }
        }
    }
    free(off); free(st);
}

static double col_dot(const double* a, const double* b, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* Multi-RHS PCG.
 * rule 0 = runPCG_multi (PCGsolver.cpp:1324-1444): stop when max_j ||r_j||/||b_j|| < tol
 *          (columns with ||b_j|| <= 1e-12 count as 0), optional stall exit after 10
 *          iterations improving the max ratio by less than 5e-3 of its previous value.
 * rule 1 = runPCG_multi_VC (varCompEst.cpp:453-528): stop when every column has
 *          ||r_j|| < tol ||b_j||.
 * rule 2 = per-column masking: a column freezes once ||r_j|| < tol ||b_j|| (or
 *          ||b_j|| <= 1e-12); the solve ends when every column is frozen.
 * x0_mode 0: x0 = 0; 1: x0 = Q^-1 b (varCompEst.cpp:467-476); 2: x0 given in x
 * (varCompEst.cpp:463-466). */
void orc_pcg(const orc_model* md, const double* qinv, const double* b, int k, double* x,
             const orc_pcg_opts* o, orc_pcg_stats* stt, double* col_ratio, int32_t* col_iters) {
    const int64_t nrec = md->nrec;
    const size_t sz = sizeof(double) * (size_t)nrec * k;
    double* r = (double*)malloc(sz);
    double* z = (double*)malloc(sz);
    double* p = (double*)malloc(sz);
    double* Vp = (double*)malloc(sz);
    double* rs = (double*)malloc(sizeof(double) * k);
    double* nb = (double*)malloc(sizeof(double) * k);
    int* active = (int*)malloc(sizeof(int) * k);
    if (o->x0_mode == 0) {
        memset(x, 0, sz);
        memcpy(r, b, sz);
    } else {
        if (o->x0_mode == 1) orc_precond_apply(md, qinv, b, k, x);
        orc_vx(md, x, k, Vp, 0);
        for (int64_t q = 0; q < nrec * k; ++q) r[q] = b[q] - Vp[q];
    }
    orc_precond_apply(md, qinv, r, k, z);
    memcpy(p, z, sz);
    for (int c = 0; c < k; ++c) {
        rs[c] = col_dot(r + c * nrec, z + c * nrec, nrec);
        nb[c] = sqrt(col_dot(b + c * nrec, b + c * nrec, nrec));
        active[c] = 1;
        if (col_iters) col_iters[c] = 0;
    }
    double last = 0.0;
    for (int c = 0; c < k; ++c) {
        double nr = sqrt(col_dot(r + c * nrec, r + c * nrec, nrec));
        if (nb[c] > 1e-12 && nr / nb[c] > last) last = nr / nb[c];
        if (o->rule == 2 && (nb[c] <= 1e-12 || nr < o->tol * nb[c])) active[c] = 0;
    }
    int stag = 0, it_done = o->max_iter, stalled = 0;
    double maxr = last;
    int any_active = 0;
    for (int c = 0; c < k; ++c) any_active |= active[c];
    if (o->rule == 2 && !any_active) it_done = 0;
    for (int it = 0; it < o->max_iter && !(o->rule == 2 && !any_active); ++it) {
        orc_vx(md, p, k, Vp, 0);
        for (int c = 0; c < k; ++c) {
            if (!active[c]) continue;
            double a = rs[c] / col_dot(p + c * nrec, Vp + c * nrec, nrec);
            for (int64_t q = 0; q < nrec; ++q) {
                x[c * nrec + q] += a * p[c * nrec + q];
                r[c * nrec + q] -= a * Vp[c * nrec + q];
            }
        }
        int done = 0;
        if (o->rule == 0) {
            maxr = 0.0;
            for (int c = 0; c < k; ++c) {
                double nr = sqrt(col_dot(r + c * nrec, r + c * nrec, nrec));
                double ratio = nb[c] > 1e-12 ? nr / nb[c] : 0.0;
                if (ratio > maxr) maxr = ratio;
            }
            if (maxr < o->tol) done = 1;
            else if (o->stagnation) {
                double impr = last - maxr, req = 5e-3 * last;
                if (impr < req) stag++; else stag = 0;
                last = maxr;
                if (stag >= 10) { done = 1; stalled = 1; }
            } else {
                last = maxr;
            }
        } else if (o->rule == 1) {
            int all = 1;
            maxr = 0.0;
            for (int c = 0; c < k; ++c) {
                double nr = sqrt(col_dot(r + c * nrec, r + c * nrec, nrec));
                if (!(nr < o->tol * nb[c])) all = 0;
                double ratio = nb[c] > 0 ? nr / nb[c] : 0.0;
                if (ratio > maxr) maxr = ratio;
            }
            done = all;
        } else {
            any_active = 0;
            maxr = 0.0;
            for (int c = 0; c < k; ++c) {
                double nr = sqrt(col_dot(r + c * nrec, r + c * nrec, nrec));
                double ratio = nb[c] > 1e-12 ? nr / nb[c] : 0.0;
                if (ratio > maxr) maxr = ratio;
                if (active[c]) {
                    if (col_iters) col_iters[c] = it + 1;
                    if (nr < o->tol * nb[c]) active[c] = 0;
                }
                any_active |= active[c];
            }
            done = !any_active;
        }
        if (done) { it_done = it + 1; break; }
        orc_precond_apply(md, qinv, r, k, z);
        for (int c = 0; c < k; ++c) {
            if (!active[c]) continue;
            double rsn = col_dot(r + c * nrec, z + c * nrec, nrec);
            double be = rsn / rs[c];
            rs[c] = rsn;
            for (int64_t q = 0; q < nrec; ++q) p[c * nrec + q] = z[c * nrec + q] + be * p[c * nrec + q];
        }
    }
    if (o->rule != 2 && col_iters)
        for (int c = 0; c < k; ++c) col_iters[c] = it_done;
    if (col_ratio)
        for (int c = 0; c < k; ++c) {
            double nr = sqrt(col_dot(r + c * nrec, r + c * nrec, nrec));
            col_ratio[c] = nb[c] > 0 ? nr / nb[c] : 0.0;
        }
    if (stt) { stt->iters = it_done; stt->stalled = stalled; stt->max_ratio = maxr; }
    free(r); free(z); free(p); free(Vp); free(rs); free(nb); free(active);
}

/* PCGsolver.cpp:1475-1513 (compute_bhat), deterministic order:
 * bhat[2i] += v_r (sig2b0 + sigb01 t_r), bhat[2i+1] += v_r (sigb01 + sig2b1 t_r). */
void orc_bhat(const orc_model* md, const double* vinv_r, double* bhat) {
    memset(bhat, 0, sizeof(double) * 2 * (size_t)md->n);
    const double sb0 = md->theta[2], sb01 = md->theta[3], sb1 = md->theta[4];
    for (int64_t r = 0; r < md->nrec; ++r) {
        int32_t i = md->indiv[r];
        double t = md->times[r];
        bhat[2 * i] += vinv_r[r] * (sb0 + sb01 * t);
        bhat[2 * i + 1] += vinv_r[r] * (sb01 + sb1 * t);
    }
}

/* PCGsolver.cpp:262-302 / varCompEst.cpp:587-685 (fitMME): solve V [X | y], then
 * beta = (X^T V^-1 X)^-1 X^T V^-1 y, V^-1 r = V^-1 y - V^-1 X beta,
 * uhat = sig2g G J^T V^-1 r (multGRM_uhat, :1447-1473), bhat (:1475-1513). */
void orc_gls_blup(const orc_model* md, const double* qinv, const double* y,
                  const orc_pcg_opts* o, double* beta, double* vinv_r, double* uhat,
                  double* bhat, double* vinv_x, double* xtvx_inv, orc_pcg_stats* st) {
    const int64_t nrec = md->nrec;
    const int q = md->p + 1, k = md->p + 2;
    double* XY = (double*)malloc(sizeof(double) * (size_t)nrec * k);
    double* S = (double*)malloc(sizeof(double) * (size_t)nrec * k);
    memcpy(XY, md->X, sizeof(double) * (size_t)nrec * q);
    memcpy(XY + (int64_t)q * nrec, y, sizeof(double) * (size_t)nrec);
    orc_pcg(md, qinv, XY, k, S, o, st, NULL, NULL);
    double* A = (double*)malloc(sizeof(double) * q * q);
    double* Ai = (double*)malloc(sizeof(double) * q * q);
    double* xty = (double*)malloc(sizeof(double) * q);
    for (int a = 0; a < q; ++a) {
        for (int b2 = 0; b2 < q; ++b2) A[a * q + b2] = col_dot(md->X + (int64_t)a * nrec, S + (int64_t)b2 * nrec, nrec);
        xty[a] = col_dot(md->X + (int64_t)a * nrec, S + (int64_t)q * nrec, nrec);
    }
    orc_inverse(q, A, Ai);
    for (int a = 0; a < q; ++a) {
        double s = 0.0;
        for (int b2 = 0; b2 < q; ++b2) s += Ai[a * q + b2] * xty[b2];
        beta[a] = s;
    }
    for (int64_t r = 0; r < nrec; ++r) {
        double v = S[(int64_t)q * nrec + r];
        for (int a = 0; a < q; ++a) v -= S[(int64_t)a * nrec + r] * beta[a];
        vinv_r[r] = v;
    }
    if (uhat) orc_vx(md, vinv_r, 1, uhat, 1);
    if (bhat) orc_bhat(md, vinv_r, bhat);
    if (vinv_x) memcpy(vinv_x, S, sizeof(double) * (size_t)nrec * q);
    if (xtvx_inv) memcpy(xtvx_inv, Ai, sizeof(double) * q * q);
    free(XY); free(S); free(A); free(Ai); free(xty);
}

/* varCompEst.cpp:693-760 (compute_f_theta_i): the five columns of dV/dtheta P y,
 * f [5][nrec]: f0 = Py; f1 = J uhat / sig2g; f2 = J J^T Py;
 * f3 = t (J^T Py)_i + (J^T (t Py))_i; f4 = t (J^T (t Py))_i. */
void orc_f_theta(const orc_model* md, const double* py, const double* uhat, double* f) {
    const int64_t nrec = md->nrec, n = md->n;
    double* w = (double*)calloc((size_t)n, sizeof(double));
    double* tw = (double*)calloc((size_t)n, sizeof(double));
    for (int64_t r = 0; r < nrec; ++r) {
        w[md->indiv[r]] += py[r];
        tw[md->indiv[r]] += py[r] * md->times[r];
    }
    for (int64_t r = 0; r < nrec; ++r) {
        int32_t i = md->indiv[r];
        double t = md->times[r];
        f[0 * nrec + r] = py[r];
        f[1 * nrec + r] = uhat[i] / md->theta[1];
        f[2 * nrec + r] = w[i];
        f[3 * nrec + r] = t * w[i] + tw[i];
        f[4 * nrec + r] = tw[i] * t;
    }
    free(w); free(tw);
}

/* varCompEst.cpp:837-862: tilde-V^-1 y by explicit block inverses (genetic part
 * replaced by sig2g * 1 within an individual). */
void orc_tildeV_solve(const orc_model* md, const double* yv, int k, double* out) {
    double* qi = (double*)malloc(sizeof(double) * (size_t)(orc_precond_size(md) + 1));
    orc_precond_build(md, NULL, qi, 1);
    orc_precond_apply(md, qi, yv, k, out);
    free(qi);
}

/* varCompEst.cpp:775-813 (sample_Yvs + compute_chol_2x2): one draw of
 * y ~ N(0, tilde-V) from standard normals z1 [n], z2 [2n], z3 [nrec]. */
void orc_sample_tildeV(const orc_model* md, const double* z1, const double* z2,
                       const double* z3, double* yv) {
    const double se = md->theta[0], sg = md->theta[1], sb0 = md->theta[2], sb01 = md->theta[3],
                 sb1 = md->theta[4];
    double l00 = sqrt(sb0), l10 = sb01 / l00, d11 = sb1 - sb01 * sb01 / sb0;
    if (d11 < 0) d11 = 0;
    double l11 = sqrt(d11);
    for (int64_t r = 0; r < md->nrec; ++r) {
        int32_t i = md->indiv[r];
        double t = md->times[r];
        yv[r] = sqrt(sg) * z1[i] + sqrt(se) * z3[r] + z2[2 * i] * (l00 + l10 * t) + z2[2 * i + 1] * l11 * t;
    }
}

/* varCompEst.cpp:980-1041 (computeDerivative_first, estimateTrace_PdV,
 * computeDerivative_AImat) at theta = md->theta with injected probes:
 *   fitMME (x0 = Q^-1 b, all-columns rule) -> Py, uhat, f_theta;
 *   first_j = Py . f_j;
 *   trace_j = mean_s (dV_j P probe_s) . weight_s, with V^-1 probe from PCG started at
 *             probe_x0 (reference: tilde-V^-1 probe) or Q^-1 probe when probe_x0 is NULL;
 *   grad = (first - trace) / 2;  AI = f^T P f / 2 with P f from PCG(x0 = Q^-1 f). */
void orc_aireml_eval(const orc_model* md, const double* y, const double* probes,
                     const double* weights, const double* probe_x0, int nmc,
                     const orc_pcg_opts* fit_opts, int nvc, double* grad, double* ai,
                     double* first_terms, double* trace) {
    const int64_t nrec = md->nrec, n = md->n;
    const int q = md->p + 1;
    double* gdiag = (double*)malloc(sizeof(double) * (size_t)n);
    orc_genetic_diag(md, gdiag);
    double* qinv = (double*)malloc(sizeof(double) * (size_t)(orc_precond_size(md) + 1));
    orc_precond_build(md, gdiag, qinv, 0);
    double beta[64];
    double* py = (double*)malloc(sizeof(double) * (size_t)nrec);
    double* uh = (double*)malloc(sizeof(double) * (size_t)n);
    double* vx = (double*)malloc(sizeof(double) * (size_t)nrec * q);
    double* xi = (double*)malloc(sizeof(double) * q * q);
    orc_pcg_opts o = *fit_opts;
    o.x0_mode = 1;
    orc_gls_blup(md, qinv, y, &o, beta, py, uh, NULL, vx, xi, NULL);
    double* f = (double*)malloc(sizeof(double) * (size_t)nrec * 5);
    orc_f_theta(md, py, uh, f);
    for (int j = 0; j < 5; ++j) first_terms[j] = col_dot(py, f + (int64_t)j * nrec, nrec);
    /* P applied to a record-space vector v (given V^-1 v): V^-1 v - V^-1 X Xi X^T V^-1 v */
    double* sv = (double*)malloc(sizeof(double) * (size_t)nrec * (nmc > 5 ? nmc : 5));
    if (nmc > 0) {
        orc_pcg_opts ot = o;
        if (probe_x0) {
            memcpy(sv, probe_x0, sizeof(double) * (size_t)nrec * nmc);
            ot.x0_mode = 2;
        } else {
            ot.x0_mode = 1;
        }
        orc_pcg(md, qinv, probes, nmc, sv, &ot, NULL, NULL, NULL);
    }
    double* pv = (double*)malloc(sizeof(double) * (size_t)nrec);
    double* us = (double*)malloc(sizeof(double) * (size_t)n);
    double* fs = (double*)malloc(sizeof(double) * (size_t)nrec * 5);
    double xv[64], cvec[64];
    for (int j = 0; j < 5; ++j) trace[j] = 0.0;
    for (int s = 0; s < nmc; ++s) {
        const double* vs = sv + (int64_t)s * nrec;
        for (int a = 0; a < q; ++a) xv[a] = col_dot(md->X + (int64_t)a * nrec, vs, nrec);
        for (int a = 0; a < q; ++a) {
            double acc = 0.0;
            for (int b2 = 0; b2 < q; ++b2) acc += xi[a * q + b2] * xv[b2];
            cvec[a] = acc;
        }
        for (int64_t r = 0; r < nrec; ++r) {
            double v = vs[r];
            for (int a = 0; a < q; ++a) v -= vx[(int64_t)a * nrec + r] * cvec[a];
            pv[r] = v;
        }
        orc_vx(md, pv, 1, us, 1);
        orc_f_theta(md, pv, us, fs);
        for (int j = 0; j < 5; ++j) trace[j] += col_dot(fs + (int64_t)j * nrec, weights + (int64_t)s * nrec, nrec);
    }
    for (int j = 0; j < 5; ++j) {
        if (nmc > 0) trace[j] /= (double)nmc;
        grad[j] = 0.5 * (first_terms[j] - trace[j]);
    }
    /* AI (varCompEst.cpp:1015-1041) over the model's nvc columns of f: 5, or 2 for the
     * reduced model (varCompEst_reducedModel_noTime.cpp:677-705), whose slope columns
     * are identically zero and would never meet the all-columns rule */
    orc_pcg_opts oa = o;
    oa.x0_mode = 1;
    orc_pcg(md, qinv, f, nvc, sv, &oa, NULL, NULL, NULL);
    for (int j = 0; j < nvc; ++j) {
        const double* vs = sv + (int64_t)j * nrec;
        for (int a = 0; a < q; ++a) xv[a] = col_dot(md->X + (int64_t)a * nrec, vs, nrec);
        for (int a = 0; a < q; ++a) {
            double acc = 0.0;
            for (int b2 = 0; b2 < q; ++b2) acc += xi[a * q + b2] * xv[b2];
            cvec[a] = acc;
        }
        for (int64_t r = 0; r < nrec; ++r) {
            double v = vs[r];
            for (int a = 0; a < q; ++a) v -= vx[(int64_t)a * nrec + r] * cvec[a];
            pv[r] = v;
        }
        for (int i2 = 0; i2 < nvc; ++i2) ai[i2 * nvc + j] = 0.5 * col_dot(f + (int64_t)i2 * nrec, pv, nrec);
    }
    free(gdiag); free(qinv); free(py); free(uh); free(vx); free(xi); free(f);
    free(sv); free(pv); free(us); free(fs);
}
