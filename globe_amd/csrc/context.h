// The context behind the C ABI: cohort, resident genetics, variance components,
// workspaces and the device PCG state of one solve stream.
#pragma once
#include <array>
#include <map>
#include <mutex>

#include "globe_internal.h"

struct ServeState;  // serving loop (serve.cpp)

struct globe_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::mutex err_mu;  // globe_serve_submit may fail on a second thread

    // cohort
    bool has_cohort = false;
    int64_t nrec = 0;
    int32_t n = 0, p = 0;
    std::vector<int32_t> h_indiv, h_blk_ind;
    std::vector<int64_t> h_blk_off;
    std::vector<double> h_times, h_X;
    globe::DevBuf d_indiv, d_times, d_X, d_blk_off, d_blk_ind;
    globe::DevBuf d_tile_blk, d_rec_blk;  // record tiles
    globe::RecMeta meta;                  // the cohort's record-space structure

    // genetics: 0 none, 1 packed genotypes, 2 GRM; gen_n = the N they were staged for
    int gen = 0;
    int32_t gen_n = 0;
    int64_t m_total = 0, m = 0, pitch_s = 0, pitch_i = 0;
    globe::DevBuf d_snp, d_ind, d_freq, d_twof, d_inv, d_grm;
    std::vector<int64_t> qc_alt, qc_miss;
    std::vector<uint8_t> qc_keep;
    std::vector<float> qc_maf, h_freq;
    globe::DevBuf d_gdiag;
    globe::DevBuf w_grm_part;    // GRM product strip partials
    globe::DevBuf w_grm_sym;     // few-column GRM product: row and transposed-tile partials
    // int8 digit tiles of the GRM for GLOBE_PREC_MIXED products, built on first use
    globe::DevBuf d_grm8, d_grm8_aux, w_grm8_part;
    bool grm8_ok = false;
    bool gdiag_ok = false;
    // genotype ingest: bytes of .bed rows per pipeline chunk, and the last ingest's
    // timings (seconds: total, host fill, host waiting on the device; chunks; threads)
    int64_t ingest_chunk = 256ll << 20;
    double ingest_stats[5] = {0, 0, 0, 0, 0};

    // int8 digits of GLOBE_PREC_MIXED contractions, and the count used outside a PCG
    // solve (globe_set_mixed_digits); while a PCG solve or the serving loop runs,
    // limbs_col holds each column's own count (relaxed on its residual), and
    // limbs_mixed their maximum
    int limbs_mixed = 4;
    int limbs_default = 4;
    std::vector<int8_t> limbs_col;

    // variance components
    std::array<double, 5> theta{0, 0, 0, 0, 0};
    bool has_theta = false;
    globe::DevBuf d_theta;       // [kMaxCols][5]
    int theta_k = -1;            // columns of d_theta filled with `theta`

    // operator workspaces
    globe::DevBuf w_S0, w_S1, w_GZ, w_P1, w_U, w_P2, w_T, w_C, w_part, w_part2, w_part3, w_part4;
    globe::DevBuf w_mix;         // mixed-precision contraction scratch
    // tuned contraction variant per shape key: block shape, partial-sum slots and (thin
    // stream-K) grid size
    struct Tuned {
        int cfg, parts;
        int64_t grid;
    };
    std::map<std::string, Tuned> tuned;
    // validation of the candidates: >= 0 runs candidate force_variant % count of every
    // contraction instead of the tuned one; variant_count = candidates of the last
    // contraction of pass 1 and pass 2
    int force_variant = -1;
    int variant_count[2] = {0, 0};
    int cu_count = 0;  // compute units of the device (stream-K grids)

    // PCG iterations replayed from captured HIP graphs (globe_set_graphs; off by default:
    // a replay runs no faster than the eager stream on ROCm 7.2)
    bool graphs_on = false;
    // Individuals a caller's genotype products involve (the serving loop: its requests'
    // individuals; S0 is zero elsewhere), gather_n entries (0: all); thin contractions
    // run pass 1 over their rows of the individual-major layout only and form only their
    // rows of G S0. globe_set_row_gather switches it off.
    const int32_t* gather_idx = nullptr;
    int64_t gather_n = 0;
    bool gather_on = true;

    // PCG workspaces and state
    globe::DevBuf p_state, p_r, p_z, p_p, p_vp;
    globe::PcgState* h_state = nullptr;  // pinned mirror

    // kernel timing: event pairs around contraction launches, tagged with the PCG
    // iteration that issued them (-1 outside PCG); a launch replayed from a graph keeps
    // its measured time instead (a = b = null)
    struct ProfEvent {
        hipEvent_t a, b;
        int pass, iter;
        double bytes, flops, issued;
        float ms = 0.f;
    };
    bool prof_on = false;
    int prof_iter = -1;
    std::vector<ProfEvent> prof;

    // host-pointer API staging
    globe::DevBuf s_in, s_out, s_aux;

    // last fitMME (reused by the AI matrix of an accepted AIREML step)
    std::vector<double> fit_xtvx_inv;
    std::array<double, 5> fit_theta{0, 0, 0, 0, 0};
    bool fit_valid = false;

    // estimator workspaces
    globe::DevBuf e_xy, e_sol, e_f, e_probe, e_w, e_sv, e_pv, e_us, e_fs, e_coef, e_small;
    // y of the AI-REML fit in progress (globe_aireml_fit holds y fixed for the whole fit
    // and uploads it once); yfit_host is the caller's array it came from, null otherwise
    globe::DevBuf e_yfit;
    const double* yfit_host = nullptr;
    // log-likelihood: probes, their solutions and the solves' Lanczos coefficients
    globe::DevBuf l_b, l_x, l_hist;

    // serving loop, when open
    ServeState* serve = nullptr;

    // No-allocation calls (globe_vx_dev_noalloc): the device's static scratch, whose
    // address hipGetSymbolAddress returned at globe_create; while `arena.base` is set,
    // every workspace of the call is a slice of it (globe::ws) and nothing is allocated.
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    struct Arena {
        uint8_t* base = nullptr;
        size_t cap = 0, used = 0;
        struct Slot {
            const void* key;
            void* p;
            size_t bytes;
        };
        Slot slot[32];
        int nslot = 0;
    } arena;
};

// Body of an ABI entry point: HIP errors and globe::Error become status codes with
// the message kept in the context.
#define API_BEGIN(ctx)                                 \
    if (!(ctx)) return GLOBE_E_INVALID;                \
    try {                                              \
        HIP_CHECK(hipSetDevice((ctx)->device));
#define API_END(ctx)                                   \
    return GLOBE_OK;                                   \
    }                                                  \
    catch (const globe::Error& e) {                    \
        std::lock_guard<std::mutex> lk((ctx)->err_mu); \
        (ctx)->err = e.what();                         \
        return e.code;                                 \
    }                                                  \
    catch (const std::exception& e) {                  \
        std::lock_guard<std::mutex> lk((ctx)->err_mu); \
        (ctx)->err = e.what();                         \
        return GLOBE_E_INVALID;                        \
    }

namespace globe {

// Workspace `b` of at least `bytes` for the current call: the context's grow-only buffer,
// or, inside a no-allocation call, a 256-byte-aligned slice of the static scratch (the
// same slice each time the call asks for the same workspace).
void* ws_bytes(globe_ctx* c, DevBuf& b, size_t bytes);
template <class T> T* ws(globe_ctx* c, DevBuf& b, size_t bytes) { return static_cast<T*>(ws_bytes(c, b, bytes)); }

void serve_release(globe_ctx* c);
// device address and size of the no-allocation scratch on `device` (the current device)
void* scratch_address(int device, size_t* bytes);

// device operator: y = V x (mode FULL) or sig2g G J^T x (GRM_ONLY); optional fused
// partials of sum x.y per column; skip = device done flag (nullptr: always run). Inside
// PCG (pst set), x is the search direction p, first updated in place to z + beta p on the
// active columns (pz = z, same leading dimension).
void op_vx(globe_ctx* c, const double* x, int64_t ldx, int k, double* y, int64_t ldy, int mode,
           int precision, const int* skip, double* part_dot, const double* pz = nullptr,
           PcgState* pst = nullptr);
void ensure_theta(globe_ctx* c, int k);
// GZ (n x k) = sig2g_c G S0 per column with the per-column components theta_cols [k][5]
void genetic_product(globe_ctx* c, const double* S0, int k, double* GZ, int precision,
                     const int* skip, const double* theta_cols);
void ensure_gdiag(globe_ctx* c);
// hist (device, (1 + 2 hist_cap) kMaxCols doubles): the Lanczos coefficients of the
// solve, laid out as PcgState::hist describes (nullptr: not kept)
void run_pcg(globe_ctx* c, const double* B, int64_t ldb, int k, double* X, int64_t ldx,
             const globe_pcg_opts& o, globe_pcg_stats* st, double* col_ratio, int32_t* col_iters,
             double* hist = nullptr, int hist_cap = 0);
// int8 digits (2..4) a mixed-precision product may use in a PCG iteration that starts at
// residual ratio |r| / |b| (inexact Krylov relaxation)
int relaxed_limbs(double ratio, double tol);
// packed-genotype contraction on the int8 matrix cores (k_mfma.hip): GZ = sig2g G S0
// with col_limbs[c] int8 digits (1..7) for column c of the right-hand side, at most 128
// digit columns in all (U_given: pass 1 already done by the caller, u = U_given, m x k;
// S0 unused; then only rows row0 .. n - 1 of GZ are formed). ridx (device, ng entries,
// ng a multiple of 512, S0 zero outside them): when the contraction is thin (at most 16
// digit columns), pass 1 contracts over those individuals only and only rows
// ridx[0 .. ng) of GZ are formed (the other rows keep what they held)
void mfma_geno_product(globe_ctx* c, const double* S0, int64_t ldn, int k, double* GZ,
                       int64_t ldg, const int* skip, const int8_t* col_limbs, const double* theta_cols,
                       const double* U_given = nullptr, int64_t row0 = 0, const int32_t* ridx = nullptr,
                       int64_t ng = 0);
// tune the packed-genotype contraction for every digit-column width a block of up to
// max_cols columns reaches at `precision` (theta_cols: [max_cols][5] components)
void pretune_geno_product(globe_ctx* c, int max_cols, int precision, const double* theta_cols);
// Rows i >= row0 of G [e_j0 .. e_j0+k-1] (n x k, row0 <= j0) from the resident genotypes
// (GRM builder: the lower triangle needs rows i >= j only)
void geno_unit_columns(globe_ctx* c, int64_t j0, int k, double* GZ, int64_t row0);
// Probes generated on the device (k_probe.hip): kind 0 Gaussian tilde-V, 1 Rademacher,
// 2 exact unit probes; seed = the Philox key.
struct ProbeGen {
    int kind;
    uint64_t seed;
};
void make_probes_dev(globe_ctx* c, int kind, int nmc, uint64_t seed, double* Y, double* Wt);
// AIREML derivative terms at the current variance components for nvc = 5 (full) or
// 2 (reduced) components, with host probes (probes, weights, probe_x0) or, when gen is
// set, probes generated on the device; y == nullptr reuses the previous fit at the same
// components (for the AI matrix of an accepted step)
int aireml_eval_impl(globe_ctx* c, const double* y, const double* probes, const double* weights,
                     const double* probe_x0, int nmc, const globe_pcg_opts* opts, bool want_ai,
                     double* grad, double* ai, double* first_terms, double* trace, int nvc,
                     const ProbeGen* gen);
// fitMME (PCGsolver.cpp:262-290, varCompEst.cpp:636-665): solve V [X | y] (y_dev on the
// device) into c->e_sol, then beta, (X^T V^-1 X)^-1 and V^-1 r into vinv_r_dev
struct Fit {
    std::vector<double> beta, xtvx_inv;  // (p+1), (p+1)^2 row-major
    globe_pcg_stats stats{};
};
Fit fit_gls(globe_ctx* c, const double* y_dev, const globe_pcg_opts& o, double* vinv_r_dev);
constexpr int kLimbsFP64 = 7;
constexpr int kLimbsMixed = 4;

}  // namespace globe
