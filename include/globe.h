/*
 * globe.h -- C ABI of libglobe_amd.so, the MI355X (gfx950) matrix-free LGMM solve.
 *
 * Plain C types only: host buffers are caller-owned and copied on entry; the
 * *_dev entry points take device pointers (allocated with globe_dev_alloc or any
 * HIP allocation on the context's device). Matrices are column-major with an
 * explicit leading dimension (the reference keeps Eigen column-major float
 * matrices); record-space matrices have nrec rows, individual-space ones n rows.
 *
 * One context per host thread (not re-entrant; the one exception is globe_serve_submit,
 * see the serving loop); all device work of a context is ordered on one HIP stream. Every call returns a globe_status; on failure
 * globe_last_error() describes it. Nothing here exits the process (the
 * reference calls exit(1) on bad input; the CLIs built on this ABI keep that).
 *
 * Each declaration cites the reference interface it replaces (ahmadazim/GLOBE).
 */
#ifndef GLOBE_H
#define GLOBE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define GLOBE_ABI_VERSION 1

typedef struct globe_ctx globe_ctx;

typedef enum {
    GLOBE_OK = 0,
    GLOBE_E_INVALID = 1,  /* bad argument, shape or cohort layout */
    GLOBE_E_STATE = 2,    /* call order (no cohort / genetics / variance components) */
    GLOBE_E_DEVICE = 3,   /* HIP runtime error or no usable device */
    GLOBE_E_NOMEM = 4,    /* device allocation failed */
    GLOBE_E_NUMERIC = 5,  /* singular system */
    GLOBE_E_IO = 6        /* input file missing or unreadable */
} globe_status;

/* globe_vx modes */
enum { GLOBE_VX_FULL = 0, GLOBE_VX_GRM_ONLY = 1 };
/* Arithmetic of the genetic product; everything else is FP64.
 * Packed genotypes, W (W^T z):
 *   FP64:      int8 matrix cores on 7 int8 digits of a 54-bit block-fixed-point z,
 *              exact int32 accumulation -- as accurate as an FP64 dot product;
 *   MIXED:     the same with 4 digits (30-bit fixed point);
 *   FP64_VALU: FP64 fused multiply-adds on the vector ALUs (validation path).
 * Dense GRM, G S0:
 *   FP64, FP64_VALU: the FP32 entries on the FP64 matrix cores, FP64 accumulation;
 *   MIXED:     int8 matrix cores on 4 int8 digits of G (per-row block fixed point of
 *              the off-diagonal entries; the diagonal applied in FP64) and of S0, exact
 *              int32 accumulation. The digit tiles (4 N^2 bytes of device memory) are
 *              built on the first mixed product after the GRM is loaded. */
enum { GLOBE_PREC_FP64 = 0, GLOBE_PREC_MIXED = 1, GLOBE_PREC_FP64_VALU = 2 };
/* PCG stopping rules */
enum {
    GLOBE_RULE_MAX_RATIO = 0,   /* runPCG_multi: max_j |r_j|/|b_j| < tol (+ optional stall exit) */
    GLOBE_RULE_ALL_COLUMNS = 1, /* runPCG_multi_VC: every |r_j| < tol |b_j| */
    GLOBE_RULE_PER_COLUMN = 2   /* per-column masking: converged columns freeze */
};
/* PCG starting points */
enum { GLOBE_X0_ZERO = 0, GLOBE_X0_PRECOND = 1, GLOBE_X0_GIVEN = 2 };

typedef struct {
    double tol;         /* reference: 1e-5 hard-coded (PCGsolver.cpp:54) */
    int32_t max_iter;   /* reference: 1000 (PCGsolver.cpp:55) */
    int32_t x0_mode;    /* GLOBE_X0_* */
    int32_t rule;       /* GLOBE_RULE_* */
    int32_t stagnation; /* rule 0 only: 10-iteration / 5e-3 stall exit (PCGsolver.cpp:1374-1415) */
    int32_t precision;  /* GLOBE_PREC_* */
    int32_t flags;      /* GLOBE_PCG_* bits, 0 by default */
} globe_pcg_opts;

/* GLOBE_PREC_MIXED solves relax the digits of the genetic product as the residual falls
 * (4, then 3, then 2 int8 digits: inexact Krylov relaxation, the operator error of
 * iteration j at most 0.01 tol |b| / |r_j|); this flag keeps 4 digits throughout. */
#define GLOBE_PCG_FIXED_DIGITS 1

typedef struct {
    int32_t iters;      /* iterations run */
    int32_t stalled;    /* 1 if the stall exit ended the solve */
    double max_ratio;   /* max_j |r_j| / |b_j| at exit */
    double seconds;     /* wall time of the solve on the device stream */
} globe_pcg_stats;

/* ---- context ------------------------------------------------------------------ */
/* No reference counterpart (the reference is one host process without a device). */
int globe_create(int device, globe_ctx** out);
void globe_destroy(globe_ctx* ctx);
/* The last failure's message on this context, valid on the calling thread until its next
 * globe_last_error call. */
const char* globe_last_error(const globe_ctx* ctx);
int globe_abi_version(void);
int globe_device_count(void);

/* ---- cohort ------------------------------------------------------------------- */
/* Replaces readInputData (PCGsolver.cpp:529-587) + block detection
 * (computePreconditioner_Block, :1004-1011). indiv: 0-based individual per record,
 * records of one individual contiguous, every index 0..n-1 present (the reference
 * silently mis-computes otherwise; here it is GLOBE_E_INVALID). X: (p+1) columns,
 * TIME last, leading dimension ldx >= nrec. */
int globe_set_cohort(globe_ctx* ctx, int64_t nrec, int32_t n, int32_t p,
                     const int32_t* indiv, const double* times, const double* X, int64_t ldx);

/* ---- genetics ------------------------------------------------------------------ */
/* Replaces prepGenotypeInput (PCGsolver.cpp:615-763): SNP-major PLINK rows
 * (m rows of `pitch` >= ceil(n/4) bytes, the .bed body after its 3 magic bytes) are
 * staged to HBM, counted on the device (bit-exact A1 / missing counts), filtered with
 * the reference's FP32 MAF / missing-rate rule and kept resident in two layouts. */
int globe_load_bed(globe_ctx* ctx, const uint8_t* rows, int64_t m, int64_t pitch,
                   float min_maf, float max_miss, int64_t* m_kept);
/* Replaces prepGenotypeInput(geno_prefix, ...) (PCGsolver.cpp:615-763) at file level:
 * M = lines of <prefix>.bim, <prefix>.bed checked (magic 6C 1B 01, at least
 * 3 + M ceil(N/4) bytes) and streamed to the device through pinned double-buffered
 * chunks (parallel pread, async copies) with the QC counts, keep rule and compaction
 * on the device. When <prefix>.fam exists its line count must equal N (the reference
 * never reads it). *m_total = M, *m_kept = SNPs kept (either may be NULL).
 * GLOBE_E_IO when a file is missing or unreadable. */
int globe_load_bed_file(globe_ctx* ctx, const char* prefix, float min_maf, float max_miss,
                        int64_t* m_total, int64_t* m_kept);
/* Per input SNP of the last globe_load_bed / globe_synth_bed: A1 allele count, missing
 * count, keep flag and FP32 A1 frequency (PCGsolver.cpp:711-734). Any pointer may be NULL. */
int globe_get_qc(globe_ctx* ctx, int64_t* alt, int64_t* miss, uint8_t* keep, float* maf);
/* Kept-SNP allele frequencies (float, as GenotypeData::allele_freqs, PCGsolver.cpp:64). */
int globe_get_freq(globe_ctx* ctx, float* freq);
/* Replaces readFullGRM (PCGsolver.cpp:778-809) / readGRM (varCompEst.cpp:270-295):
 * n(n+1)/2 float32 lower triangle, row-major (GCTA .grm.bin). */
int globe_load_grm(globe_ctx* ctx, const float* grm_lt);
/* GRM builder (the reference only declares one, PCGsolver.cpp:165-167): G = W W^T / M
 * from the resident kept genotypes (W standardised as in getStdGenotype,
 * PCGsolver.cpp:831-842), computed exactly on the int8 matrix cores column block by
 * column block and stored as the FP32 packed lower triangle of a .grm.bin; the context
 * then works from the GRM as after globe_load_grm (the genotype layouts are released). */
int globe_grm_from_geno(globe_ctx* ctx);
/* The context's GRM as the n(n+1)/2 FP32 lower triangle, row-major (GCTA .grm.bin order). */
int globe_get_grm(globe_ctx* ctx, float* grm_lt);
/* Rows r0 .. r1 - 1 of that triangle (r1(r1+1)/2 - r0(r0+1)/2 floats, starting with
 * element (r0, 0)), so a writer can stream a GRM larger than host memory (80 GB at
 * N = 200k). The reference declares a GRM builder but ships none (PCGsolver.cpp:165-167). */
int globe_get_grm_rows(globe_ctx* ctx, int64_t r0, int64_t r1, float* rows);
/* Replaces computeGeneticDiagonal (PCGsolver.cpp:925-989): diag(G), n doubles. */
int globe_genetic_diag(globe_ctx* ctx, double* out);

/* ---- variance components ------------------------------------------------------------- */
/* Replaces SolvingData's sig2e..sig2b1 (PCGsolver.cpp:47-52) and the rebuild of the
 * block preconditioner (computePreconditioner_Block :991-1058, computePreconditioner
 * varCompEst.cpp:355-415). theta = {sig2e, sig2g, sig2b0, sigb01, sig2b1}. */
int globe_set_varcomp(globe_ctx* ctx, const double theta[5]);

/* ---- operator ----------------------------------------------------------------------- */
/* Replaces computeVx_multi (PCGsolver.cpp:1095-1266) and computeVx_multi_VC
 * (varCompEst.cpp:322-352). mode GLOBE_VX_FULL: Y (nrec x k) = V X;
 * GLOBE_VX_GRM_ONLY (multGRM_uhat): Y (n x k) = sig2g G J^T X. */
int globe_vx(globe_ctx* ctx, const double* X, int64_t ldx, int k, double* Y, int64_t ldy,
             int mode, int precision);
int globe_vx_dev(globe_ctx* ctx, const double* X, int64_t ldx, int k, double* Y, int64_t ldy,
                 int mode, int precision);
/* No counterpart in the reference: globe_vx_dev for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same operator, modes and precisions, same kernels, results bit-identical to
 * globe_vx_dev; every workspace is a slice of one static 128 MiB device array that each
 * device the library runs on holds until the process ends, so nothing is allocated and
 * nothing is tuned (the contraction runs its candidate with the fewest partial slots).
 * The work is enqueued on the context stream and the call returns without waiting
 * (globe_synchronize waits). GLOBE_E_INVALID when the call needs more than the scratch
 * (24 N k bytes of individual-space vectors, 8 M k of SNP-space ones, the contraction's
 * digit matrix and split partials: N = 20k, M = 40k fits up to k = 32, N = 200k,
 * M = 500k fits k = 1 in FP64); GLOBE_E_STATE when the call would have to allocate
 * (mixed-precision GRM digits not built yet). The array is shared by every context on
 * a device and keeps the last call's contents: calls that overlap on one device race. */
int globe_vx_dev_noalloc(globe_ctx* ctx, const double* X, int64_t ldx, int k, double* Y, int64_t ldy,
                         int mode, int precision);
/* Replaces applyPreconditioner (PCGsolver.cpp:1060-1093, varCompEst.cpp:432-450):
 * Z = Q^-1 R per individual block, all k columns. */
int globe_precond_apply(globe_ctx* ctx, const double* R, int64_t ldr, int k, double* Z, int64_t ldz);

/* ---- solver --------------------------------------------------------------------------- */
/* Replaces runPCG_multi (PCGsolver.cpp:1324-1444) and runPCG_multi_VC
 * (varCompEst.cpp:453-528). Solves V X = B for k columns; X holds x0 on entry when
 * x0_mode == GLOBE_X0_GIVEN. col_ratio (k) and col_iters (k) may be NULL. */
int globe_pcg(globe_ctx* ctx, const double* B, int64_t ldb, int k, double* X, int64_t ldx,
              const globe_pcg_opts* opts, globe_pcg_stats* stats, double* col_ratio,
              int32_t* col_iters);
int globe_pcg_dev(globe_ctx* ctx, const double* B, int64_t ldb, int k, double* X, int64_t ldx,
                  const globe_pcg_opts* opts, globe_pcg_stats* stats, double* col_ratio,
                  int32_t* col_iters);

/* ---- estimators ---------------------------------------------------------------------- */
/* Replaces runPCGsolver's main body (PCGsolver.cpp:262-302): PCG on [X | y], GLS
 * beta (p+1), V^-1 r (nrec), uhat = sig2g G J^T V^-1 r (n, compute_uhat :1447-1473),
 * bhat (2n interleaved [b0, b1], compute_bhat :1475-1513). Output pointers may be NULL. */
int globe_gls_blup(globe_ctx* ctx, const double* y, const globe_pcg_opts* opts, double* beta,
                   double* vinv_r, double* uhat, double* bhat, globe_pcg_stats* stats);
/* Replaces computeDerivative_first + estimateTrace_PdV + computeDerivative_AImat
 * (varCompEst.cpp:980-1041) at the current variance components: fitMME (x0 = Q^-1 b,
 * all-columns rule), f_theta, trace of P dV_theta from nmc injected probes
 * (probes nrec x nmc; weights nrec x nmc: tilde-V^-1 probe for the reference's
 * Gaussian tilde-V probes, the probe itself for Rademacher probes; probe_x0 nrec x nmc
 * PCG starting point or NULL for Q^-1 probe), gradient (5) and AI matrix (5x5,
 * row-major). first_terms / trace (5 each) may be NULL. */
int globe_aireml_eval(globe_ctx* ctx, const double* y, const double* probes,
                      const double* weights, const double* probe_x0, int nmc,
                      const globe_pcg_opts* opts, double* grad, double* ai,
                      double* first_terms, double* trace);

/* Replaces the NLopt LD_MMA trust-region step of oneIter_AIREML (varCompEst.cpp:1162-1214,
 * objective :1044-1065, constraints :1068-1145): minimise -(g.p - p^T AI p / 2) over
 * -10 <= p_i <= 10 subject to sum_i AI_ii p_i <= delta, theta_i + p_i >= 1e-5 (i != 3)
 * and (theta_3+p_3)^2 - (theta_2+p_2)(theta_4+p_4) + 1e-3 <= 0. Returns the Newton step
 * AI^-1 g when it is feasible, else the best constrained minimiser of a multi-start
 * augmented-Lagrangian projected-Newton solve. Host only (no context). GLOBE_E_NUMERIC when no feasible
 * step is found (the reference then shrinks the trust region, :1201-1214). */
int globe_aireml_step(const double grad[5], const double ai[25], const double theta[5],
                      double delta, double step[5]);

/* ---- reduced model (varCompEst_reducedModel_noTime.cpp): V = sig2g A + sig2e I ----------
 * One record per individual (globe_set_cohort with nrec == n; records in GRM order), no
 * time term: the covariates are passed as X with all their columns (no TIME column) and
 * the context's components are (sig2e, sig2g, 0, 0, 0). On that cohort the per-individual
 * block preconditioner is the reference's diagonal 1 / (sig2g A_ii + sig2e)
 * (computePreconditioner :326-335). */
/* Replaces computeDerivative_first + estimateTrace_PdV + computeDerivative_AImat of the
 * reduced model (varCompEst_reducedModel_noTime.cpp:592-705): fitMME (:440-518), f_theta
 * = [V^-1 r, A V^-1 r] (:524-548), the trace of P dV_theta from nmc injected probes (as
 * globe_aireml_eval), gradient (2) and AI matrix (2x2 row-major). GLOBE_E_INVALID unless
 * nrec == n and sig2b0 = sigb01 = sig2b1 = 0. */
int globe_aireml_eval_reduced(globe_ctx* ctx, const double* y, const double* probes,
                              const double* weights, const double* probe_x0, int nmc,
                              const globe_pcg_opts* opts, double* grad, double* ai,
                              double* first_terms, double* trace);
/* Replaces the NLopt LD_MMA step of the reduced model's oneIter_AIREML
 * (varCompEst_reducedModel_noTime.cpp:784-840, objective :707-728, constraints
 * :731-781): minimise -(g.p - p^T AI p / 2) over -10 <= p_i <= 10 subject to
 * sum_i AI_ii p_i <= delta and theta_i + p_i >= 1e-5 (no covariance constraint). Same
 * solver and return codes as globe_aireml_step. */
int globe_aireml_step_reduced(const double grad[2], const double ai[4], const double theta[2],
                              double delta, double step[2]);

/* Replaces generate_standard_normal + sample_Yvs and the tilde-V^-1 solve of
 * estimateTrace_PdV (varCompEst.cpp:763-862): nmc probes (nrec x nmc, column-major) and
 * their trace weights at the context's variance components, generated on the device
 * from Philox4x32-10 keyed by `seed`. kind 0: Gaussian tilde-V probes (weights tilde-V^-1
 * probe; the reference draws them from std::random_device); 1: Rademacher (weights = the
 * probes); 2: unit vectors sqrt(nrec) e_s (nmc == nrec; the trace terms are then exact).
 * Either output may be NULL. globe_aireml_fit uses seed globe_probe_seed(opts.seed, e)
 * for its evaluation e = 0, 1, ...; these calls return the same probes, for
 * globe_aireml_eval. */
int globe_aireml_probes(globe_ctx* ctx, int kind, int nmc, uint64_t seed, double* probes,
                        double* weights);
uint64_t globe_probe_seed(uint64_t seed, uint64_t eval_id);

typedef void (*globe_log_fn)(const char* line, void* user);

typedef struct {
    int32_t max_iter;   /* --maxiter, reference default 50 (varCompEst.cpp:21) */
    double tol;         /* --tol, reference default 1e-3 (:22) */
    int32_t num_mc;     /* --numMC, reference default 15 (:23) */
    int32_t probes;     /* 0: Gaussian tilde-V probes (reference, :775-797); 1: Rademacher;
                           2: exact -- unit probes sqrt(nrec) e_s, numMC must equal nrec
                           (small cohorts: the trace terms are then exact) */
    uint64_t seed;      /* probe and restart RNG seed (the reference's is irreproducible) */
    globe_pcg_opts pcg; /* inner solves; reference: tol 1e-5, max 1000, all-columns rule */
} globe_aireml_opts;

/* Replaces runAIREML's main loop + oneIter_AIREML (varCompEst.cpp:1148-1491): AI-REML with
 * the reference's trust-region control (step halving, rho test, Delta update, restarts,
 * Delta < 1e-4 stop). theta: start values in, estimates out (sig2e, sig2g, sig2b0,
 * sigb01, sig2b1); se = sqrt(diag(AI^-1)); ai 5x5 row-major; iters and mean seconds per
 * iteration out; log (may be NULL) receives the progress lines the reference prints. */
int globe_aireml_fit(globe_ctx* ctx, const double* y, const globe_aireml_opts* opts,
                     double theta[5], double se[5], double ai[25], int32_t* iters,
                     double* sec_per_iter, globe_log_fn log, void* user);
/* Replaces the reduced model's main loop + oneIter_AIREML
 * (varCompEst_reducedModel_noTime.cpp:784-1099): the same trust-region control as
 * globe_aireml_fit over theta = (sig2e, sig2g) (start in, estimates out), se (2), ai
 * (2x2 row-major). The cohort must hold one record per individual. */
int globe_aireml_fit_reduced(globe_ctx* ctx, const double* y, const globe_aireml_opts* opts,
                             double theta[2], double se[2], double ai[4], int32_t* iters,
                             double* sec_per_iter, globe_log_fn log, void* user);

/* ---- REML log-likelihood (no reference counterpart: the reference's AI-REML loop tracks
 * only predicted and approximate changes of it, varCompEst.cpp:1221, 1277-1278) ---- */
typedef struct {
    double loglik;       /* -1/2 (log|V| + log|X^T V^-1 X| + y^T P y + (Nrec - p - 1) log 2pi) */
    double logdet_v;     /* log|V|: logdet_q + the Lanczos quadrature estimate */
    double logdet_v_se;  /* standard error of logdet_v over the probes (0 for exact probes) */
    double logdet_q;     /* log|Q| of the block preconditioner, exact */
    double logdet_xtvx;  /* log|X^T V^-1 X| */
    double ypy;          /* y^T P y */
    int32_t probes;      /* probe solves run */
    int32_t max_steps;   /* most Lanczos steps (PCG iterations) of one probe */
    double seconds;      /* wall time of the call */
} globe_loglik;

/* The REML log-likelihood at the context's variance components. log|V| = log|Q| +
 * log|Q^-1/2 V Q^-1/2|: log|Q| exactly over the per-individual preconditioner blocks,
 * the second term by stochastic Lanczos quadrature -- probes b = F u with F F^T = Q,
 * solved V x = b by the block PCG (per-column rule, x0 = 0, opts->tol / max_iter /
 * precision), each column's alpha/beta giving the Lanczos matrix T and the estimate
 * (b^T Q^-1 b) e1^T (log(T) - T + I) e1 (the term T - I has the known mean tr(Q^-1 V) -
 * Nrec = 0 and cuts the spread several-fold). kind 0: Gaussian u, 1: Rademacher u (nprobe >= 2 each,
 * estimate = their mean); 2: exact -- the Nrec + 2N columns of F (nprobe ignored), whose
 * sum is log|Q^-1/2 V Q^-1/2| itself up to the solves' tolerance. Probes come from
 * Philox4x32-10 keyed by `seed`. The fixed effects are solved as globe_gls_blup does.
 * Needs sig2e > 0, sig2g >= 0, sig2b0 >= 0, sig2b1 >= 0 and sigb01^2 <= sig2b0 sig2b1. */
int globe_reml_loglik(globe_ctx* ctx, const double* y, int kind, int nprobe, uint64_t seed,
                      const globe_pcg_opts* opts, globe_loglik* out);

/* ---- serving loop (no reference counterpart: the reference solves one trait per run) -- */
/* One resident cohort (globe_set_cohort for N, genotypes or GRM loaded) serves solve
 * requests V_j x_j = y_j, each with its own individuals, records, times and variance
 * components. Requests are admitted into the columns ("slots") of one shared block PCG
 * as slots and HBM free up -- continuous batching: a request joins between two
 * iterations of the running block and leaves when its own column converges. Each
 * column's arithmetic (per-column fixed-point scale of the contraction, its own PCG
 * scalars and record structure) is independent of the other columns, so a result is
 * the same alone and in the mix. While the loop is open the context runs nothing else. */
typedef struct {
    int64_t nrec;         /* records of the request */
    const int32_t* indiv; /* 0-based cohort individual per record, each individual's records contiguous */
    const double* times;  /* nrec */
    const double* y;      /* nrec: the right-hand side (a trait) */
    double theta[5];      /* sig2e, sig2g, sig2b0, sigb01, sig2b1 of this request */
} globe_request;

typedef struct {
    int32_t slots;        /* columns of the shared block PCG (1..64) */
    int64_t max_records;  /* record capacity of a slot */
    int64_t hbm_budget;   /* bytes of the record pool that holds every admitted request's
                             record structure and PCG vectors: 80 bytes per element, a
                             request takes nrec + 1 elements rounded up to 64 (0: room for
                             `slots` requests of max_records) */
    double tol;           /* per-request |r| < tol |b| */
    int32_t max_iter;     /* per request */
    int32_t precision;    /* GLOBE_PREC_* of the genotype contraction */
} globe_serve_opts;

typedef struct {
    int64_t id;
    int32_t iters;
    int32_t converged;
    double ratio;         /* |r| / |b| at exit */
    double t_submit;      /* seconds on the loop's clock (from globe_serve_open) */
    double t_admit;
    double t_done;
} globe_serve_result;

/* Requests are admitted first in, first out, each into a free slot and a first-fit range
 * of the record pool; a request's data is uploaded while the block iterates and it joins
 * at the next step. */
int globe_serve_open(globe_ctx* ctx, const globe_serve_opts* opts);
/* Copies, validates and queues a request; *id identifies it in globe_serve_take. This one
 * call may run on other host threads (request intake, one or several at once) concurrently
 * with the thread that calls globe_serve_step, so requests are prepared on the host while
 * the GPU iterates. */
int globe_serve_submit(globe_ctx* ctx, const globe_request* req, int64_t* id);
/* Admits queued requests (FIFO) into free slots while the HBM budget allows, runs one
 * iteration of the block PCG over the occupied slots and retires the requests whose
 * column converged (or reached max_iter). Counts after the step may be NULL: finished
 * (results ready this step), active (requests in the block), queued (requests waiting,
 * including those staged to join at the next step). */
int globe_serve_step(globe_ctx* ctx, int32_t* finished, int32_t* active, int32_t* queued);
/* Ids of finished requests not yet taken (up to max, ascending); *count gets how many. */
int globe_serve_finished(globe_ctx* ctx, int64_t* ids, int32_t max, int32_t* count);
/* Hands over a finished request: x (its nrec values) and its result; GLOBE_E_STATE if it
 * has not finished. Each finished request can be taken once. */
int globe_serve_take(globe_ctx* ctx, int64_t id, double* x, globe_serve_result* res);
/* Record-pool use: out[0] bytes in use, out[1] the most in use at once, out[2] the pool's
 * bytes, out[3] steps at which the queue's head waited for pool room with a slot free. */
int globe_serve_usage(globe_ctx* ctx, int64_t out[4]);
int globe_serve_close(globe_ctx* ctx);

/* ---- mixed-precision digits (no reference counterpart) ----------------------------- */
/* int8 digits (2..4, default 4) of GLOBE_PREC_MIXED genetic products applied outside a PCG
 * solve (globe_vx and friends). PCG solves start at 4 and lower the count themselves as
 * the residual falls, unless GLOBE_PCG_FIXED_DIGITS is set; they leave this value in
 * place. Fewer digits trade accuracy (about 2^(1-8L) of each operand's scale) for speed. */
int globe_set_mixed_digits(globe_ctx* ctx, int digits);

/* ---- PCG iteration graphs (no reference counterpart) ------------------------------- */
/* on: a PCG solve captures one iteration's launches into a HIP graph per digit
 * configuration (after running that configuration once eagerly) and replays it; off (the
 * default): every iteration's kernels are enqueued one by one. Results are identical. */
int globe_set_graphs(globe_ctx* ctx, int on);

/* ---- serving: rows of the genotype product (no reference counterpart) -------------- */
/* on (the default): while at most 80% of the cohort's individuals belong to the serving
 * loop's requests and the contraction is thin (at most 16 digit columns), both genotype
 * passes read only their rows of the individual-major layout: pass 1 contracts over those
 * individuals only (S0 is zero elsewhere), pass 2 forms only their rows of G S0 (the
 * other rows are never read); off: every individual and row. Results are identical. */
int globe_set_row_gather(globe_ctx* ctx, int on);

/* ---- contraction variants (no reference counterpart: validation of the autotuner) ---- */
/* The first packed-genotype contraction of each shape times every candidate variant (block
 * shape, split-K or stream-K grid) and keeps the fastest; all candidates compute the same
 * exact integer sums. index >= 0: every later contraction runs candidate index % count of
 * its shape's candidate list instead, untimed; -1 (the default) returns to the tuned
 * variants. globe_variant_counts: candidates of pass 1 and pass 2 of the last contraction. */
int globe_force_variant(globe_ctx* ctx, int index);
int globe_variant_counts(globe_ctx* ctx, int32_t counts[2]);

/* ---- kernel timing (no reference counterpart) --------------------------------------- */
/* When enabled, every genetic-product launch is bracketed by HIP events on the context
 * stream (launches that a finished PCG skips are discarded). which: 0 = pass 1 of the
 * packed-genotype contraction (SNP-major rows, W^T z), 1 = pass 2 (individual-major rows,
 * W u), 2 = dense GRM product in FP64, 3 = dense GRM product on int8 digits (mixed).
 * bytes / flops are the algorithmic figures per launch: rows x ceil(K/4) packed bytes and
 * 2 N M k flops for the genotypes; 4 N(N+1)/2 bytes (the FP32 triangle once) and 2 N^2 k
 * flops for the GRM. */
typedef struct {
    int64_t launches;
    double ms_total;
    double bytes_per_launch;
    double flops_per_launch;      /* averaged over the launches */
    double issued_ops_per_launch; /* int8 MFMA operations actually issued (digits, padding),
                                     averaged over the launches */
} globe_kernel_stats;
int globe_profile_enable(globe_ctx* ctx, int on);
int globe_profile_read(globe_ctx* ctx, int which, globe_kernel_stats* out);

/* ---- device memory and synthetic cohorts (no reference counterpart) ---------------- */
/* Device buffers for the *_dev entry points and for measurement with inputs resident in HBM. */
int globe_dev_alloc(globe_ctx* ctx, int64_t bytes, void** ptr);
/* Free and total device memory of the context's device, bytes (e.g. to size a serving
 * loop's HBM budget). */
int globe_device_memory(globe_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);
int globe_dev_free(globe_ctx* ctx, void* ptr);
int globe_memcpy_h2d(globe_ctx* ctx, void* dst, const void* src, int64_t bytes);
int globe_memcpy_d2h(globe_ctx* ctx, void* dst, const void* src, int64_t bytes);
int globe_synchronize(globe_ctx* ctx);
/* Bytes of .bed rows per ingest pipeline chunk (default 256 MiB); tests use small
 * chunks to cross chunk boundaries. */
int globe_set_ingest_chunk(globe_ctx* ctx, int64_t bytes);
/* The last host-sourced genotype ingest: seconds in total, seconds filling pinned
 * buffers on the host, seconds the host waited for a buffer, chunks, host threads. */
int globe_ingest_stats(globe_ctx* ctx, double out[5]);
/* Plain device-to-device copy kernel (16-byte vector loads / stores); the reference
 * point for HBM byte counters (bytes a multiple of 16). */
int globe_dev_copy(globe_ctx* ctx, void* dst, const void* src, int64_t bytes);
/* Generates an m-SNP x n-individual PLINK cohort directly in HBM (counter-based hash
 * RNG, seeded): f_s ~ U(f_lo, f_hi), dosage ~ Binomial(2, f_s), missing w.p. miss;
 * then runs the same QC and staging as globe_load_bed. For cohorts too large to
 * write to disk; the generator never runs inside a timed region. */
int globe_synth_bed(globe_ctx* ctx, int64_t m, uint64_t seed, float f_lo, float f_hi,
                    float miss, float min_maf, float max_miss, int64_t* m_kept);
/* out (n) = sum_s w_is a_s: standardized kept genotypes times an SNP-space vector
 * (the synthetic phenotype recipe's u = sqrt(sig2g / M) W a). */
int globe_geno_Wa(globe_ctx* ctx, const double* a, double* out);
/* Rows of the resident SNP-major layout (pitch bytes each) for checks, written back
 * in PLINK codes: dosage 2 -> 00, 1 -> 10, 0 -> 11. A missing call (01) held as
 * dosage 0 therefore reads back as 11, and padding reads as 11. */
int globe_get_geno_rows(globe_ctx* ctx, int64_t first, int64_t count, uint8_t* out, int64_t* pitch);

#ifdef __cplusplus
}
#endif
#endif
