/*
 * globe_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's matrix-free longitudinal genomic mixed-model
 * (LGMM) solve, written from the reference's math (ahmadazim/GLOBE, src/PCGsolver.cpp,
 * src/varCompEst.cpp). It is the checker for the HIP product path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The product
 * (globe_amd/, libglobe_amd.so) never links or calls it.
 *
 * Arithmetic: FP64 everywhere except
 *   - orc_bed_qc(), which reproduces the reference's FP32 MAF / missing-rate
 *     decision bit for bit (PCGsolver.cpp:726-732), and
 *   - orc_vx_ref32(), an FP32 OpenMP restatement of computeVx_multi kept only as
 *     the CPU timing baseline (same per-element decode and 50,000-SNP chunking).
 *
 * Layouts: record-space matrices are column-major [k][nrec] (Eigen's layout in
 * the reference); individual-space matrices are column-major [k][n].
 */
#ifndef GLOBE_ORACLE_H
#define GLOBE_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Model state: cohort + genetics + variance components. */
typedef struct {
    int64_t nrec;            /* records */
    int32_t n;               /* individuals (unique IDs) */
    int32_t p;               /* covariates excluding TIME; X has p+1 columns */
    const int32_t* indiv;    /* [nrec] 0-based individual index per record */
    const double* times;     /* [nrec] */
    const double* X;         /* [(p+1)][nrec] col-major, TIME in the last column */
    int32_t use_grm;         /* 0: packed genotypes, 1: GRM lower triangle */
    const uint8_t* geno;     /* [m][pitch] kept SNP rows, SNP-major PLINK 2-bit */
    int64_t m;               /* kept SNPs */
    int64_t pitch;           /* bytes per SNP row (>= ceil(n/4)) */
    const float* freq;       /* [m] A1 allele frequency as stored by the reference (float) */
    const float* grm_lt;     /* [n(n+1)/2] FP32 lower triangle, row-major (GCTA .grm.bin) */
    double theta[5];         /* sig2e, sig2g, sig2b0, sigb01, sig2b1 */
} orc_model;

/* --- genotype decode / QC ------------------------------------------------- */
int orc_decode(uint8_t code);
void orc_bed_counts(const uint8_t* rows, int64_t m, int32_t n, int64_t pitch,
                    int64_t* alt, int64_t* miss);
int64_t orc_bed_qc(const uint8_t* rows, int64_t m, int32_t n, int64_t pitch,
                   float min_maf, float max_miss, uint8_t* keep, float* maf_out,
                   float* miss_rate_out);
double orc_inv_std(float f);

/* --- operator --------------------------------------------------------------- */
void orc_genetic_diag(const orc_model* md, double* gdiag);
void orc_vx(const orc_model* md, const double* x, int k, double* y, int grm_only);
void orc_vx_dense(const orc_model* md, const double* x, int k, double* y);
void orc_geno_Wa(const orc_model* md, const double* a, double* out);
void orc_vx_ref32(const orc_model* md, const float* x, int k, float* y, int grm_only);
void orc_vx_vc_ref32(const orc_model* md, const float* x, int k, float* y);

/* --- preconditioner ----------------------------------------------------------- */
int64_t orc_blocks(const orc_model* md, int64_t* starts /* nrec+1 */);
int64_t orc_precond_size(const orc_model* md);
void orc_precond_build(const orc_model* md, const double* gdiag, double* qinv, int no_grm);
void orc_precond_apply(const orc_model* md, const double* qinv, const double* r, int k, double* z);

/* --- PCG ------------------------------------------------------------------------ */
typedef struct {
    double tol;
    int32_t max_iter;
    int32_t x0_mode;   /* 0 zero, 1 Q^-1 b, 2 caller-provided X */
    int32_t rule;      /* 0 max-ratio (+stagnation), 1 all-columns, 2 per-column masking */
    int32_t stagnation;/* rule 0 only: 1 enables the 10-iteration / 5e-3 stall exit */
} orc_pcg_opts;

typedef struct {
    int32_t iters;
    int32_t stalled;
    double max_ratio;
} orc_pcg_stats;

void orc_pcg(const orc_model* md, const double* qinv, const double* b, int k, double* x,
             const orc_pcg_opts* o, orc_pcg_stats* st, double* col_ratio, int32_t* col_iters);

/* --- estimators ------------------------------------------------------------------------ */
int orc_inverse(int n, const double* a, double* ainv);
void orc_gls_blup(const orc_model* md, const double* qinv, const double* y,
                  const orc_pcg_opts* o, double* beta, double* vinv_r, double* uhat,
                  double* bhat, double* vinv_x /* [(p+1)][nrec] or NULL */,
                  double* xtvx_inv /* [(p+1)^2] or NULL */, orc_pcg_stats* st);
void orc_bhat(const orc_model* md, const double* vinv_r, double* bhat);
void orc_f_theta(const orc_model* md, const double* py, const double* uhat, double* f);
void orc_tildeV_solve(const orc_model* md, const double* yv, int k, double* out);
void orc_sample_tildeV(const orc_model* md, const double* z1, const double* z2,
                       const double* z3, double* yv);
void orc_aireml_eval(const orc_model* md, const double* y, const double* probes,
                     const double* weights, const double* probe_x0, int nmc,
                     const orc_pcg_opts* fit_opts, int nvc, double* grad, double* ai,
                     double* first_terms, double* trace);

#ifdef __cplusplus
}
#endif
#endif
