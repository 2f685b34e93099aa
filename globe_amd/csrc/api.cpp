// libglobe_amd.so: the C ABI (include/globe.h) over the HIP kernels.
//
// Host code owns orchestration only: staging, the operator sequence, the PCG loop
// (whose scalars live on the device), the small FP64 dense algebra of GLS and
// AIREML ((p+1)^2 and 5x5 systems), and error reporting. Every vector operation
// runs on the device stream of the context.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <list>
#include <numeric>

#include "context.h"

using globe::DevBuf;
using globe::Error;

namespace globe {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw Error(e == hipErrorOutOfMemory ? GLOBE_E_NOMEM : GLOBE_E_DEVICE,
                    std::string(what) + ": " + hipGetErrorString(e));
}

// set while a no-allocation entry point runs on this thread: any device allocation is
// then a bug, reported instead of performed
thread_local bool t_noalloc = false;
void set_noalloc(bool on) { t_noalloc = on; }
// set while an iteration is captured into a HIP graph: an allocation there would be a
// workspace the graph does not own (every workspace is sized by an eager run first)
thread_local bool t_capturing = false;
// moves on every device allocation and release: captured graphs hold workspace addresses
// and are discarded once it has moved
std::atomic<uint64_t> g_alloc_epoch{0};

void DevBuf::ensure(size_t b) {
    if (b <= bytes && p) return;
    require(!t_noalloc, GLOBE_E_STATE, "device allocation inside a no-allocation call");
    require(!t_capturing, GLOBE_E_STATE, "device allocation inside a captured PCG iteration");
    release();
    ++g_alloc_epoch;
    if (b == 0) b = 256;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
        p = nullptr;
        bytes = 0;
        (void)hipGetLastError();
        throw Error(GLOBE_E_NOMEM, "device allocation of " + std::to_string(b) + " bytes failed: " +
                                       hipGetErrorString(e));
    }
    bytes = b;
}

void DevBuf::release() {
    if (p) {
        (void)hipFree(p);
        ++g_alloc_epoch;
    }
    p = nullptr;
    bytes = 0;
}

void* ws_bytes(globe_ctx* c, DevBuf& b, size_t bytes) {
    globe_ctx::Arena& a = c->arena;
    if (!a.base) {
        b.ensure(bytes);
        return b.p;
    }
    for (int q = 0; q < a.nslot; ++q)
        if (a.slot[q].key == &b && a.slot[q].bytes >= bytes) return a.slot[q].p;
    const size_t off = (size_t)round_up((int64_t)a.used, 256);
    require(off + bytes <= a.cap && a.nslot < 32, GLOBE_E_INVALID,
            "the call needs more than the " + std::to_string(a.cap >> 20) +
                " MiB of static scratch of the no-allocation entry point (use the allocating one)");
    void* p = a.base + off;
    a.used = off + bytes;
    a.slot[a.nslot++] = {&b, p, bytes};
    return p;
}

// Gauss-Jordan inverse with partial pivoting (the reference inverts X^T V^-1 X and
// the AI matrix with Eigen's inverse(), PCGsolver.cpp:280, varCompEst.cpp:1485).
static bool invert(int n, const std::vector<double>& a, std::vector<double>& out) {
    std::vector<double> w((size_t)n * 2 * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            w[i * 2 * n + j] = a[i * n + j];
            w[i * 2 * n + n + j] = i == j ? 1.0 : 0.0;
        }
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(w[r * 2 * n + c]) > std::fabs(w[piv * 2 * n + c])) piv = r;
        if (w[piv * 2 * n + c] == 0.0) return false;
        if (piv != c)
            for (int j = 0; j < 2 * n; ++j) std::swap(w[c * 2 * n + j], w[piv * 2 * n + j]);
        const double d = w[c * 2 * n + c];
        for (int j = 0; j < 2 * n; ++j) w[c * 2 * n + j] /= d;
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = w[r * 2 * n + c];
            if (f == 0.0) continue;
            for (int j = 0; j < 2 * n; ++j) w[r * 2 * n + j] -= f * w[c * 2 * n + j];
        }
    }
    out.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) out[i * n + j] = w[i * 2 * n + n + j];
    return true;
}

void ensure_theta(globe_ctx* c, int k) {
    require(c->has_theta, GLOBE_E_STATE, "variance components not set (globe_set_varcomp)");
    if (c->theta_k >= k) return;
    c->d_theta.ensure(sizeof(double) * 5 * kMaxCols);
    std::vector<double> h((size_t)5 * kMaxCols);
    for (int j = 0; j < kMaxCols; ++j)
        for (int q = 0; q < 5; ++q) h[j * 5 + q] = c->theta[q];
    HIP_CHECK(hipMemcpyAsync(c->d_theta.p, h.data(), h.size() * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->theta_k = kMaxCols;
}

void ensure_gdiag(globe_ctx* c) {
    if (c->gdiag_ok) return;
    require(c->gen != 0, GLOBE_E_STATE, "no genetics loaded (globe_load_bed / globe_load_grm)");
    c->d_gdiag.ensure(sizeof(double) * (size_t)c->n);
    if (c->gen == 1) {
        launch_gdiag(c->stream, c->d_ind.as<uint8_t>(), c->pitch_i, c->n, c->m,
                     c->d_twof.as<double>(), c->d_inv.as<double>(), c->d_gdiag.as<double>());
    }
    // GRM: the diagonal was uploaded by globe_load_grm
    c->gdiag_ok = true;
}

// int8 digit tiles of the GRM (4 N^2 bytes: the full square) for mixed products
static void ensure_grm_digits(globe_ctx* c) {
    if (c->grm8_ok) return;
    const int32_t n = c->n;
    c->d_grm8.ensure(grm_digit_bytes(n));
    c->d_grm8_aux.ensure(grm_digit_aux_bytes(n));
    launch_grm_digits(c->stream, c->d_grm.as<float>(), n, c->d_grm8_aux.p, c->d_grm8.as<int8_t>());
    c->grm8_ok = true;
}

void genetic_product(globe_ctx* c, const double* S0, int k, double* GZ, int precision,
                     const int* skip, const double* th) {
    const int32_t n = c->n;
    hipStream_t s = c->stream;
    if (c->gen == 2) {
        const bool mixed = precision == GLOBE_PREC_MIXED;
        // FP64: up to 8 columns read each stored tile once; wider products run passes of
        // up to 32 columns over the strips, and a last group of at most 8 columns (k = 33 to
        // 40, 65 to 72, ...) takes the read-once product instead of a strip pass
        const int tail = mixed || k <= kSymMaxCols ? 0 : (k % 32 <= kSymMaxCols ? k % 32 : 0);
        const int kstrip = mixed || k <= kSymMaxCols ? 0 : k - tail;
        const int ksym = mixed ? 0 : (k <= kSymMaxCols ? k : tail);
        void* scratch = nullptr;
        void* scratch_sym = nullptr;
        if (mixed) {
            ensure_grm_digits(c);
            ensure_gdiag(c);
            scratch = ws_bytes(c, c->w_grm8_part, grm_i8_part_bytes(n));
        } else {
            if (kstrip > 0) scratch = ws_bytes(c, c->w_grm_part, grm_product_part_bytes(n));
            if (ksym > 0) scratch_sym = ws_bytes(c, c->w_grm_sym, grm_sym_part_bytes(n, ksym));
        }
        // timing: which 2 = FP64 product, 3 = int8 digit product; algorithmic bytes = the
        // FP32 triangle once, flops = 2 N^2 k
        globe_ctx::ProfEvent e{};
        const bool prof = c->prof_on && !c->arena.base;
        if (prof) {
            const double np = (double)((n + 63) / 64 * 64), L = c->limbs_mixed;
            const double cols = 16.0 * ((k + 15) / 16);
            HIP_CHECK(hipEventCreate(&e.a));
            HIP_CHECK(hipEventCreate(&e.b));
            e.pass = mixed ? 3 : 2;
            e.iter = c->prof_iter;
            e.bytes = 4.0 * (double)n * (n + 1) / 2;
            e.flops = 2.0 * (double)n * n * k;
            const int symc = ksym <= 2 ? ksym : ksym <= 4 ? 4 : 8;
            e.issued = mixed ? 2.0 * np * np * cols * (L * (L + 1) / 2 + L - 1)
                             : 2.0 * np * np * ((kstrip > 0 ? grm_padded_cols(kstrip) : 0) + symc);
            // external: inside a captured PCG iteration the record is a graph node that
            // stamps the event on every replay
            HIP_CHECK(hipEventRecordWithFlags(e.a, s, hipEventRecordExternal));
        }
        if (mixed)
            launch_grm_product_i8(s, c->d_grm8.as<int8_t>(), grm_digit_rexp(c->d_grm8_aux.p, n), n, S0, n, k,
                                  c->limbs_mixed, c->d_gdiag.as<double>(), th, GZ, n, skip, scratch);
        else {
            if (kstrip > 0)
                launch_grm_product(s, c->d_grm.as<float>(), n, S0, n, kstrip, th, GZ, n, skip,
                                   static_cast<double*>(scratch));
            if (ksym > 0)
                launch_grm_product_sym(s, c->d_grm.as<float>(), n, S0 + (size_t)kstrip * n, n, ksym,
                                       th + 5 * kstrip, GZ + (size_t)kstrip * n, n, skip,
                                       static_cast<double*>(scratch_sym));
        }
        if (prof) {
            HIP_CHECK(hipEventRecordWithFlags(e.b, s, hipEventRecordExternal));
            c->prof.push_back(e);
        }
        return;
    }
    if (c->m == 0) {
        launch_fill_cols(s, n, GZ, n, k, 0.0);
        return;
    }
    if (precision != GLOBE_PREC_FP64_VALU) {
        // digits per column: 7 for FP64; mixed: each column's own relaxed count while a
        // solve sets them, else the uniform count
        std::vector<int8_t> L((size_t)k);
        const bool per_col = precision == GLOBE_PREC_MIXED && (int)c->limbs_col.size() >= k;
        for (int q = 0; q < k; ++q)
            L[q] = precision != GLOBE_PREC_MIXED ? (int8_t)kLimbsFP64
                   : per_col                     ? c->limbs_col[q]
                                                 : (int8_t)c->limbs_mixed;
        // column groups whose digits fit one contraction (128 digit columns)
        for (int c0 = 0; c0 < k;) {
            int c1 = c0, J = 0;
            while (c1 < k && J + L[c1] <= 128) J += L[c1++];
            mfma_geno_product(c, S0 + (size_t)c0 * n, n, c1 - c0, GZ + (size_t)c0 * n, n, skip, L.data() + c0,
                              th + (size_t)c0 * 5, nullptr, 0, c->gather_n > 0 ? c->gather_idx : nullptr,
                              c->gather_n);
            c0 = c1;
        }
        return;
    }
    const int64_t m = c->m;
    double* P1 = ws<double>(c, c->w_P1, sizeof(double) * (size_t)m * k);
    double* U = ws<double>(c, c->w_U, sizeof(double) * (size_t)m * k);
    double* P2 = ws<double>(c, c->w_P2, sizeof(double) * (size_t)n * k);
    double* T = ws<double>(c, c->w_T, sizeof(double) * kMaxCols);
    double* Cs = ws<double>(c, c->w_C, sizeof(double) * kMaxCols);
    double* part = ws<double>(c, c->w_part2, sizeof(double) * kMaxParts * kMaxCols);
    // pass 1 (PCGsolver.cpp:1143-1171): P1[c][s] = sum_i g_is S0[c][i]
    launch_rowdot(s, c->d_snp.as<uint8_t>(), c->pitch_s, m, n, S0, n, k, P1, m, skip);
    launch_colsum_parts(s, n, S0, n, k, part, skip);
    launch_reduce_parts_skip(s, part, parts_for(n), k, T, skip);
    launch_center_snp(s, m, k, P1, m, T, c->d_twof.as<double>(), c->d_inv.as<double>(), th, 1.0 / (double)m, U, m,
                      part, skip);
    launch_reduce_parts_skip(s, part, parts_for(m), k, Cs, skip);
    // pass 2 (PCGsolver.cpp:1173-1199): GZ[c][i] = sum_s g_is u[c][s] - C[c]
    launch_rowdot(s, c->d_ind.as<uint8_t>(), c->pitch_i, n, m, U, m, k, P2, n, skip);
    launch_center_ind(s, n, k, P2, n, Cs, GZ, n, skip);
}

void op_vx(globe_ctx* c, const double* x, int64_t ldx, int k, double* y, int64_t ldy, int mode,
           int precision, const int* skip, double* part_dot, const double* pz, PcgState* pst) {
    require(c->has_cohort, GLOBE_E_STATE, "cohort not set (globe_set_cohort)");
    require(c->gen != 0, GLOBE_E_STATE, "no genetics loaded (globe_load_bed / globe_load_grm)");
    require(k >= 1 && k <= kMaxCols, GLOBE_E_INVALID, "k must be in [1, 64]");
    ensure_theta(c, k);
    const int32_t n = c->n;
    double* S0 = ws<double>(c, c->w_S0, sizeof(double) * (size_t)n * k);
    double* S1 = ws<double>(c, c->w_S1, sizeof(double) * (size_t)n * k);
    double* GZ = ws<double>(c, c->w_GZ, sizeof(double) * (size_t)n * k);
    hipStream_t s = c->stream;
    if (pst) {
        // x is PCG's search direction: p = z + beta p first (fused into the sums when the
        // records are tiled)
        if (!launch_seg_sums_pcg(s, c->meta, const_cast<double*>(x), pz, ldx, k, pst, S0, S1, n)) {
            launch_pcg_p(s, pst, c->nrec, const_cast<double*>(x), pz, ldx, k);
            launch_seg_sums(s, c->meta, x, ldx, k, S0, S1, n, skip);
        }
    } else {
        launch_seg_sums(s, c->meta, x, ldx, k, S0, S1, n, skip);
    }
    if (mode == GLOBE_VX_GRM_ONLY) {
        genetic_product(c, S0, k, GZ, precision, skip, c->d_theta.as<double>());
        launch_copy_cols(s, n, GZ, n, y, ldy, k);
        return;
    }
    genetic_product(c, S0, k, GZ, precision, skip, c->d_theta.as<double>());
    launch_record_out(s, c->meta, x, ldx, k, GZ, S0, S1, n, c->d_theta.as<double>(), y, ldy, part_dot,
                      skip);
}

static void precond(globe_ctx* c, const double* r, int64_t ldr, int k, double* z, int64_t ldz,
                    double* part_rz, double* part_rr, const PcgState* st) {
    ensure_gdiag(c);
    ensure_theta(c, k);
    launch_precond(c->stream, c->meta, c->d_gdiag.as<double>(), c->d_theta.as<double>(), r, ldr, k, z,
                   ldz, part_rz, part_rr, st, 0);
}

// nmc probes (nrec x nmc) and their trace weights at the context's variance components:
// Gaussian tilde-V probes (kind 0) are weighted by tilde-V^-1 y, the per-individual
// blocks of sig2g 1 1^T + U B U^T + sig2e I (estimateTrace_PdV, varCompEst.cpp:839-862),
// applied by the preconditioner kernel with G_ii := 1; the others by themselves.
void make_probes_dev(globe_ctx* c, int kind, int nmc, uint64_t seed, double* Y, double* Wt) {
    const int64_t nrec = c->nrec;
    hipStream_t s = c->stream;
    launch_probes(s, kind, nrec, nmc, seed, c->d_indiv.as<int32_t>(), c->d_times.as<double>(),
                  c->theta.data(), Y, nrec);
    if (kind == 0) {
        ensure_gdiag(c);
        ensure_theta(c, nmc);
        launch_precond(s, c->meta, c->d_gdiag.as<double>(), c->d_theta.as<double>(), Y, nrec, nmc, Wt, nrec,
                       nullptr, nullptr, nullptr, 1);
    } else {
        launch_copy_cols(s, nrec, Y, nrec, Wt, nrec, nmc);
    }
}

// PCG with the reference's control flow (PCGsolver.cpp:1324-1444 for rule 0,
// varCompEst.cpp:453-528 for rule 1), all scalars on the device. The host enqueues
// batches of iterations and reads back only the done flag between batches.
// Digits of the mixed contraction for a column whose iteration starts at residual ratio
// |r_j| / |b_j| (inexact Krylov relaxation: Simoncini & Szyld, SIAM J. Sci.
// Comput. 25 (2003); van den Eshof & Sleijpen, SIAM J. Matrix Anal. Appl. 26 (2004)).
// An operator error of relative size eta in iteration j moves the true residual away
// from the recursive one by about eta |r_j|, so eta_j <= 0.01 tol |b| / |r_j| keeps the
// gap below 0.01 tol |b| per iteration. L int8 digits of the block fixed point bound the
// quantisation error of a column by 2^-(8L-2) of its largest entry; 8 2^-(8L-2) is taken
// as eta(L) (the largest entry of a right-hand side is a few times its rms). Measured at
// N = 200k, M = 500k, 32 RHS, tol 1e-6: 0.01 keeps the 11 iterations of the 4-digit
// solve (0.1 costs a twelfth) with 3.3 digits on average.
int relaxed_limbs(double ratio, double tol) {
    const double allowed = 0.01 * tol / std::max(ratio, 1e-300);
    for (int L = 2; L < kLimbsMixed; ++L)
        if (8.0 * std::ldexp(1.0, -(8 * L - 2)) <= allowed) return L;
    return kLimbsMixed;
}

namespace {

// One PCG iteration's launches replayed from a captured HIP graph. Within one solve the
// launch sequence of an iteration depends only on its digit configuration (each
// column's int8 digits in a mixed solve), so a graph is kept per configuration. A
// configuration first runs eagerly (that run tunes every contraction shape it needs and
// sizes every workspace) and is captured on its next use. Captured kernel arguments are
// workspace addresses: a device allocation or release since the capture discards it.
struct IterGraphs {
    struct Entry {
        std::vector<int8_t> key;
        uint64_t warm = ~0ull;   // allocation epoch of its last eager run
        uint64_t epoch = 0;      // allocation epoch at capture
        hipGraphExec_t exec = nullptr;
        std::vector<globe_ctx::ProfEvent> prof;  // timing events recorded by the graph
    };
    std::list<Entry> e;  // stable addresses: a batch keeps pointers to the entries it launched
    ~IterGraphs() {
        for (Entry& x : e) drop(x);
    }
    static void drop(Entry& x) {
        if (x.exec) (void)hipGraphExecDestroy(x.exec);
        x.exec = nullptr;
        for (auto& p : x.prof) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        x.prof.clear();
    }
    Entry& find(const std::vector<int8_t>& key) {
        for (Entry& x : e)
            if (x.key == key) return x;
        e.push_back(Entry{});
        e.back().key = key;
        return e.back();
    }
};

}  // namespace

void run_pcg(globe_ctx* c, const double* B, int64_t ldb, int k, double* X, int64_t ldx,
             const globe_pcg_opts& o, globe_pcg_stats* st, double* col_ratio, int32_t* col_iters,
             double* hist, int hist_cap) {
    require(c->has_cohort && c->gen != 0, GLOBE_E_STATE, "cohort and genetics required");
    require(k >= 1 && k <= kMaxCols, GLOBE_E_INVALID, "k must be in [1, 64]");
    require(o.rule >= 0 && o.rule <= 2, GLOBE_E_INVALID, "unknown PCG rule");
    require(o.x0_mode >= 0 && o.x0_mode <= 2, GLOBE_E_INVALID, "unknown x0 mode");
    require(o.precision >= GLOBE_PREC_FP64 && o.precision <= GLOBE_PREC_FP64_VALU, GLOBE_E_INVALID,
            "unknown precision");
    ensure_gdiag(c);
    ensure_theta(c, k);
    const int64_t nrec = c->nrec;
    hipStream_t s = c->stream;
    const size_t vb = sizeof(double) * (size_t)nrec * k;
    c->p_r.ensure(vb);
    c->p_z.ensure(vb);
    c->p_p.ensure(vb);
    c->p_vp.ensure(vb);
    c->p_state.ensure(sizeof(PcgState));
    c->w_part.ensure(sizeof(double) * kMaxParts * kMaxCols);
    c->w_part3.ensure(sizeof(double) * kMaxParts * kMaxCols);
    c->w_part4.ensure(sizeof(double) * kMaxParts * kMaxCols);
    double* r = c->p_r.as<double>();
    double* z = c->p_z.as<double>();
    double* p = c->p_p.as<double>();
    double* vp = c->p_vp.as<double>();
    double* part_a = c->w_part.as<double>();
    double* part_b = c->w_part3.as<double>();
    double* part_c = c->w_part4.as<double>();
    PcgState* dst = c->p_state.as<PcgState>();
    PcgState* hs = c->h_state;
    std::memset(hs, 0, sizeof(PcgState));
    hs->max_iter = o.max_iter;
    hs->rule = o.rule;
    hs->stagnation = o.stagnation;
    hs->k = k;
    hs->tol = o.tol;
    hs->hist = hist;
    hs->hist_cap = hist ? hist_cap : 0;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, s));
    HIP_CHECK(hipMemcpyAsync(dst, hs, sizeof(PcgState), hipMemcpyHostToDevice, s));
    const int np = parts_for(nrec);
    const int npb = precond_parts(c->meta);
    // starting point and residual
    if (o.x0_mode == GLOBE_X0_ZERO) {
        launch_fill_cols(s, nrec, X, ldx, k, 0.0);
        launch_copy_cols(s, nrec, B, ldb, r, nrec, k);
    } else {
        if (o.x0_mode == GLOBE_X0_PRECOND) precond(c, B, ldb, k, X, ldx, nullptr, nullptr, nullptr);
        op_vx(c, X, ldx, k, vp, nrec, GLOBE_VX_FULL, o.precision, nullptr, nullptr);
        launch_sub_cols(s, nrec, B, ldb, vp, nrec, r, nrec, k);
    }
    launch_dot_parts(s, nrec, B, ldb, B, ldb, k, part_a);
    precond(c, r, nrec, k, z, nrec, part_c, part_b, nullptr);
    launch_copy_cols(s, nrec, z, nrec, p, nrec, k);
    // partial slot counts: dots over records use np, precond uses npb
    // (k_pcg_init reads up to max(np, npb) slots; unused slots are zero-filled below)
    const int nmax = std::max(np, npb);
    launch_zero_slots(s, part_b, part_c, nullptr, k, npb, nmax);
    launch_zero_slots(s, part_a, nullptr, nullptr, k, np, nmax);
    launch_pcg_init(s, dst, part_a, part_b, part_c, nmax, k);
    const int* skip = &dst->done;
    // Iterations are enqueued in batches that double while an iteration is short
    // (launch-bound problems), capped one above the iterations the residual's geometric
    // decay over the last batch predicts are left (iterations enqueued past convergence
    // return at once, but their launches still cost). A relaxed mixed solve whose product
    // costs less with fewer digits (a GRM, or 17 or more digit columns) also ends a batch
    // where that decay predicts its digits can drop, since a batch keeps the digits of its
    // start; one whose product does not (a thin genotype pass, bound by the rows it
    // streams) keeps doubling, as each batch boundary is a host round trip. When one
    // iteration takes over 2 ms the host checks the done flag after every iteration, so no
    // skipped launches are issued.
    int batch = 1, queued = 0;
    double last_ratio = 1.0;
    const bool digits_matter = c->gen == 2 || kLimbsMixed * k > 16;
    const size_t prof_mark = c->prof.size();
    auto tb = std::chrono::steady_clock::now();
    // mixed precision: digits per batch from the residual ratio at its start (4 for the
    // first iteration); a batch of several iterations only gets more digits than it needs
    const bool relax = o.precision == GLOBE_PREC_MIXED && !(o.flags & GLOBE_PCG_FIXED_DIGITS);
    struct LimbsGuard {
        globe_ctx* c;
        ~LimbsGuard() {
            c->limbs_mixed = c->limbs_default;
            c->limbs_col.clear();
        }
    } guard{c};
    c->limbs_mixed = kLimbsMixed;
    c->limbs_col.clear();
    const bool tiled = c->meta.cap > 0;
    auto iteration = [&]() {
        // p = z + beta p (beta = 0 in the first iteration, where p = z already) fused into
        // the segmented sums of V p; the dot partials come out of the kernels that produce
        // their operands (V p output; x/r update + preconditioner, one pass, which also
        // reduces alpha in every workgroup); one single-block reduction: the convergence
        // test with beta
        op_vx(c, p, nrec, k, vp, nrec, GLOBE_VX_FULL, o.precision, skip, part_a, z, dst);
        if (tiled) {
            launch_pcg_xr_precond(s, c->meta, c->d_gdiag.as<double>(), c->d_theta.as<double>(), dst, part_a, np, X,
                                  ldx, r, p, vp, nrec, k, z, part_c, part_b);
            launch_pcg_check_beta(s, dst, part_b, npb, part_c, npb, k);
        } else {
            launch_pcg_alpha(s, dst, part_a, np, k);
            launch_pcg_xr(s, dst, nrec, X, ldx, r, p, vp, nrec, k, part_b);
            precond(c, r, nrec, k, z, nrec, part_c, nullptr, dst);
            launch_pcg_check_beta(s, dst, part_b, np, part_c, npb, k);
        }
    };
    // graphs (not inside a no-allocation call, whose workspaces are arena slices); with
    // timing on, one iteration per batch, so the host reads each replay's events before
    // the next replay records them again
    const bool graphs = c->graphs_on && !c->arena.base;
    IterGraphs ig;
    std::vector<IterGraphs::Entry*> launched;
    while (true) {
        if (graphs && c->prof_on) batch = 1;
        launched.clear();
        for (int b = 0; b < batch; ++b) {
            c->prof_iter = queued + b;
            if (!graphs) {
                iteration();
                continue;
            }
            std::vector<int8_t> key(c->limbs_col);
            key.push_back((int8_t)c->limbs_mixed);
            IterGraphs::Entry& g = ig.find(key);
            const uint64_t now = g_alloc_epoch.load();
            if (g.exec && g.epoch != now) IterGraphs::drop(g);
            if (!g.exec && g.warm == now) {
                // capture: the same launches, recorded instead of run
                const size_t pm = c->prof.size();
                hipGraph_t graph = nullptr;
                HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                t_capturing = true;
                try {
                    iteration();
                } catch (...) {
                    t_capturing = false;
                    (void)hipStreamEndCapture(s, &graph);
                    if (graph) (void)hipGraphDestroy(graph);
                    for (size_t q = pm; q < c->prof.size(); ++q) {
                        (void)hipEventDestroy(c->prof[q].a);
                        (void)hipEventDestroy(c->prof[q].b);
                    }
                    c->prof.resize(pm);
                    throw;
                }
                t_capturing = false;
                HIP_CHECK(hipStreamEndCapture(s, &graph));
                const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HIP_CHECK(ie);
                g.epoch = now;
                g.prof.assign(c->prof.begin() + pm, c->prof.end());
                c->prof.resize(pm);
            }
            if (g.exec) {
                HIP_CHECK(hipGraphLaunch(g.exec, s));
                launched.push_back(&g);
            } else {
                iteration();
                g.warm = g_alloc_epoch.load();
            }
        }
        queued += batch;
        // the whole state (3 KB: the header, and the per-column ratios and flags the
        // relaxation and the caller read) after every batch, so the last batch's copy is
        // the final one; the end event goes with it
        HIP_CHECK(hipEventRecord(e1, s));
        HIP_CHECK(hipMemcpyAsync(hs, dst, sizeof(PcgState), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        // timing events of a replayed graph: read now (batch 1), before a replay re-records them
        for (IterGraphs::Entry* g : launched)
            for (const auto& e : g->prof) {
                globe_ctx::ProfEvent t = e;
                t.a = t.b = nullptr;
                t.iter = queued - 1;
                HIP_CHECK(hipEventElapsedTime(&t.ms, e.a, e.b));
                c->prof.push_back(t);
            }
        if (hs->done || queued >= o.max_iter + 1) break;
        if (relax) {
            // each column's digits from its own residual ratio (the columns are
            // independent recurrences, so each gets the single-RHS bound); a column the
            // per-column rule has frozen takes none: its V p is never read again (its x
            // and r stay, its alpha is 0)
            c->limbs_col.assign((size_t)k, (int8_t)kLimbsMixed);
            int lmax = 2, live = 0;
            for (int j = 0; j < k; ++j) {
                const bool frozen = o.rule == GLOBE_RULE_PER_COLUMN && !hs->active[j];
                const int L = frozen ? 0 : relaxed_limbs(hs->ratio[j], o.tol);
                c->limbs_col[j] = (int8_t)L;
                lmax = std::max(lmax, L);
                live += L > 0;
            }
            if (!live) c->limbs_col[0] = 2;  // the batch's launches are no-ops once done
            c->limbs_mixed = lmax;
        }
        const auto tn = std::chrono::steady_clock::now();
        const double per_iter = std::chrono::duration<double>(tn - tb).count() / batch;
        tb = tn;
        int left = 16;
        const double ratio = hs->max_ratio;
        if ((!relax || digits_matter) && ratio > 0.0 && ratio < last_ratio && ratio > o.tol) {
            const double per = std::log(ratio / last_ratio) / batch;  // log of the decay per iteration
            left = (int)std::min(16.0, std::ceil(std::log(o.tol / ratio) / per) + 1.0);
            const int L = relax ? relaxed_limbs(ratio, o.tol) : 2;
            if (L > 2) {
                // the ratio below which relaxed_limbs gives one digit fewer
                const double drop = 0.01 * o.tol / (8.0 * std::ldexp(1.0, -(8 * (L - 1) - 2)));
                left = std::min(left, std::max(1, (int)std::ceil(std::log(drop / ratio) / per)));
            }
        }
        last_ratio = ratio;
        batch = per_iter > 2e-3 ? 1 : std::max(1, std::min({batch * 2, 16, left}));
    }
    c->prof_iter = -1;
    if (c->prof_on) {
        // launches of iterations after convergence returned at once: drop them
        std::vector<globe_ctx::ProfEvent> keep(c->prof.begin(), c->prof.begin() + prof_mark);
        for (size_t q = prof_mark; q < c->prof.size(); ++q) {
            const auto& e = c->prof[q];
            if (e.iter < hs->iters) {
                keep.push_back(e);
            } else if (e.a) {
                (void)hipEventDestroy(e.a);
                (void)hipEventDestroy(e.b);
            }
        }
        c->prof.swap(keep);
    }
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) {
        st->iters = hs->iters;
        st->stalled = hs->stalled;
        st->max_ratio = hs->max_ratio;
        st->seconds = ms * 1e-3;
    }
    for (int j = 0; j < k; ++j) {
        if (col_ratio) col_ratio[j] = hs->ratio[j];
        if (col_iters) col_iters[j] = hs->col_iters[j];
    }
}

}  // namespace globe

using namespace globe;

// ----------------------------------------------------------------------------- ABI

extern "C" {

int globe_abi_version(void) { return GLOBE_ABI_VERSION; }

int globe_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int globe_create(int device, globe_ctx** out) {
    if (!out) return GLOBE_E_INVALID;
    *out = nullptr;
    int n = globe_device_count();
    if (device < 0 || device >= n) return GLOBE_E_DEVICE;
    globe_ctx* c = new globe_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipHostMalloc((void**)&c->h_state, sizeof(PcgState), hipHostMallocDefault) != hipSuccess) {
        globe_destroy(c);
        return GLOBE_E_DEVICE;
    }
    // the device address of the static scratch of the no-allocation entry point
    try {
        c->scratch = scratch_address(device, &c->scratch_bytes);
    } catch (const Error&) {
        globe_destroy(c);
        return GLOBE_E_DEVICE;
    }
    *out = c;
    return GLOBE_OK;
}

void globe_destroy(globe_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    serve_release(c);
    for (auto& e : c->prof) {
        if (!e.a) continue;
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    if (c->h_state) (void)hipHostFree(c->h_state);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// The message is copied under the context's error lock into a per-thread buffer, so a
// submit failing on an intake thread cannot free the string another thread is reading.
const char* globe_last_error(const globe_ctx* c) {
    if (!c) return "null context";
    thread_local std::string msg;
    std::lock_guard<std::mutex> lk(const_cast<globe_ctx*>(c)->err_mu);
    msg = c->err;
    return msg.c_str();
}

int globe_set_cohort(globe_ctx* c, int64_t nrec, int32_t n, int32_t p, const int32_t* indiv,
                     const double* times, const double* X, int64_t ldx) {
    API_BEGIN(c)
    require(nrec > 0 && n > 0 && p >= 0 && indiv && times && X && ldx >= nrec, GLOBE_E_INVALID,
            "globe_set_cohort: bad sizes or null pointers");
    require(c->serve == nullptr, GLOBE_E_STATE, "globe_set_cohort: close the serving loop first");
    // blocks = runs of equal individual (PCGsolver.cpp:1004-1011); each individual
    // must form exactly one run and every index 0..n-1 must occur.
    std::vector<int64_t> off;
    std::vector<int32_t> ind;
    std::vector<uint8_t> seen((size_t)n, 0);
    for (int64_t r = 0; r < nrec; ++r) {
        const int32_t i = indiv[r];
        // per-record checks: the message is built only on failure
        if (i < 0 || i >= n)
            throw Error(GLOBE_E_INVALID, "individual index " + std::to_string(i) + " at record " +
                                             std::to_string(r) + " outside [0, n)");
        if (r == 0 || indiv[r - 1] != i) {
            if (seen[i])
                throw Error(GLOBE_E_INVALID,
                            "records of individual " + std::to_string(i + 1) + " are not contiguous");
            seen[i] = 1;
            off.push_back(r);
            ind.push_back(i);
        }
    }
    off.push_back(nrec);
    require((int64_t)ind.size() == n, GLOBE_E_INVALID,
            "individual IDs must be exactly 1..N: " + std::to_string(ind.size()) +
                " distinct individuals for N = " + std::to_string(n));
    c->nrec = nrec;
    c->n = n;
    c->p = p;
    c->h_indiv.assign(indiv, indiv + nrec);
    c->h_times.assign(times, times + nrec);
    c->h_X.resize((size_t)nrec * (p + 1));
    for (int q = 0; q <= p; ++q)
        std::memcpy(&c->h_X[(size_t)q * nrec], X + (size_t)q * ldx, sizeof(double) * nrec);
    c->h_blk_off = off;
    c->h_blk_ind = ind;
    c->d_indiv.ensure(sizeof(int32_t) * nrec);
    c->d_times.ensure(sizeof(double) * nrec);
    c->d_X.ensure(sizeof(double) * nrec * (p + 1));
    c->d_blk_off.ensure(sizeof(int64_t) * (n + 1));
    c->d_blk_ind.ensure(sizeof(int32_t) * n);
    HIP_CHECK(hipMemcpy(c->d_indiv.p, indiv, sizeof(int32_t) * nrec, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->d_times.p, times, sizeof(double) * nrec, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->d_X.p, c->h_X.data(), sizeof(double) * nrec * (p + 1), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->d_blk_off.p, off.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->d_blk_ind.p, ind.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    // record tiles: greedy runs of whole blocks, at most kTileRecords records and
    // kTileBlocks blocks each
    {
        std::vector<int32_t> tile_blk{0}, rec_blk((size_t)nrec);
        int64_t maxblk = 0, in_tile = 0;
        for (int32_t b = 0; b < n; ++b) {
            const int64_t len = off[b + 1] - off[b];
            maxblk = std::max(maxblk, len);
            if (in_tile > 0 && (in_tile + len > kTileRecords || b - tile_blk.back() == kTileBlocks)) {
                tile_blk.push_back(b);
                in_tile = 0;
            }
            in_tile += len;
            for (int64_t r = off[b]; r < off[b + 1]; ++r) rec_blk[r] = b;
        }
        tile_blk.push_back(n);
        RecMeta& mt = c->meta;
        mt = RecMeta{};
        mt.indiv = c->d_indiv.as<int32_t>();
        mt.times = c->d_times.as<double>();
        mt.blk_off = c->d_blk_off.as<int64_t>();
        mt.blk_ind = c->d_blk_ind.as<int32_t>();
        mt.nrec = nrec;
        mt.nblk = n;
        if (maxblk <= kTileMaxRecords) {
            c->d_tile_blk.ensure(sizeof(int32_t) * tile_blk.size());
            c->d_rec_blk.ensure(sizeof(int32_t) * std::max<int64_t>(nrec, 1));
            HIP_CHECK(hipMemcpy(c->d_tile_blk.p, tile_blk.data(), sizeof(int32_t) * tile_blk.size(),
                                hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(c->d_rec_blk.p, rec_blk.data(), sizeof(int32_t) * nrec, hipMemcpyHostToDevice));
            mt.tile_blk = c->d_tile_blk.as<int32_t>();
            mt.rec_blk = c->d_rec_blk.as<int32_t>();
            mt.ntiles = (int32_t)tile_blk.size() - 1;
            mt.cap = kTileRecords;
        }
    }
    c->has_cohort = true;
    c->fit_valid = false;
    // Genotypes and GRMs are laid out for the N they were staged with (row pitch, row
    // count, tiles, diag(G)); a cohort of another size releases them, so the next
    // product fails with GLOBE_E_STATE instead of indexing buffers of the old size.
    if (c->gen != 0 && c->gen_n != n) {
        c->gen = 0;
        c->gen_n = 0;
        c->d_snp.release();
        c->d_ind.release();
        c->d_grm.release();
        c->d_grm8.release();
        c->d_grm8_aux.release();
        c->m = 0;
        c->grm8_ok = false;
        c->gdiag_ok = false;
        c->tuned.clear();
    }
    API_END(c)
}

}  // extern "C"

extern "C" {

int globe_get_qc(globe_ctx* c, int64_t* alt, int64_t* miss, uint8_t* keep, float* maf) {
    API_BEGIN(c)
    require(c->gen == 1, GLOBE_E_STATE, "no genotypes loaded");
    const size_t m = c->qc_alt.size();
    if (alt) std::memcpy(alt, c->qc_alt.data(), sizeof(int64_t) * m);
    if (miss) std::memcpy(miss, c->qc_miss.data(), sizeof(int64_t) * m);
    if (keep) std::memcpy(keep, c->qc_keep.data(), m);
    if (maf) std::memcpy(maf, c->qc_maf.data(), sizeof(float) * m);
    API_END(c)
}

int globe_get_freq(globe_ctx* c, float* freq) {
    API_BEGIN(c)
    require(c->gen == 1, GLOBE_E_STATE, "no genotypes loaded");
    if (freq) std::memcpy(freq, c->h_freq.data(), sizeof(float) * c->h_freq.size());
    API_END(c)
}

int globe_get_geno_rows(globe_ctx* c, int64_t first, int64_t count, uint8_t* out, int64_t* pitch) {
    API_BEGIN(c)
    require(c->gen == 1, GLOBE_E_STATE, "no genotypes loaded");
    require(first >= 0 && count >= 0 && first + count <= c->m, GLOBE_E_INVALID, "row range");
    if (pitch) *pitch = c->pitch_s;
    if (out && count > 0) {
        const size_t nb = (size_t)count * c->pitch_s;
        HIP_CHECK(hipMemcpy(out, c->d_snp.as<uint8_t>() + first * c->pitch_s, nb, hipMemcpyDeviceToHost));
        // 2-bit dosage d -> PLINK code: 2 -> 00, 1 -> 10, 0 -> 11
        uint8_t lut[256];
        for (int b = 0; b < 256; ++b) {
            int v = 0;
            for (int q = 0; q < 4; ++q) {
                const int d = (b >> (2 * q)) & 3;
                v |= (d == 2 ? 0 : d == 1 ? 2 : 3) << (2 * q);
            }
            lut[b] = (uint8_t)v;
        }
        for (size_t q = 0; q < nb; ++q) out[q] = lut[out[q]];
    }
    API_END(c)
}

namespace {

// Rows r0 .. r1 of the packed triangle per chunk, at most `cap` floats (one row at least).
int64_t grm_chunk_end(int64_t r0, int64_t n, int64_t cap) {
    auto span = [](int64_t a, int64_t b) { return (b * (b + 1) - a * (a + 1)) / 2; };
    int64_t r1 = std::min<int64_t>(n, r0 + std::max<int64_t>(1, cap / (r0 + 1)));
    while (r1 > r0 + 1 && span(r0, r1) > cap) r1 = r0 + (r1 - r0) / 2;
    return r1;
}

constexpr int64_t kGrmStageFloats = 64ll << 20;  // 256 MB of packed triangle per chunk

}  // namespace

int globe_load_grm(globe_ctx* c, const float* lt) {
    API_BEGIN(c)
    require(c->has_cohort, GLOBE_E_STATE, "cohort must be set before the GRM");
    require(lt != nullptr, GLOBE_E_INVALID, "null GRM");
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    c->gen = 0;
    c->d_snp.release();
    c->d_ind.release();
    c->m = 0;
    // the packed triangle streams through a staging buffer into the tile-major layout
    c->d_grm.ensure(grm_tile_bytes((int32_t)n));
    HIP_CHECK(hipMemsetAsync(c->d_grm.p, 0, grm_tile_bytes((int32_t)n), s));
    DevBuf stage;
    stage.ensure(sizeof(float) * (size_t)std::min<int64_t>(kGrmStageFloats, n * (n + 1) / 2 + n));
    for (int64_t r0 = 0; r0 < n;) {
        const int64_t r1 = grm_chunk_end(r0, n, kGrmStageFloats);
        const int64_t off = r0 * (r0 + 1) / 2, cnt = r1 * (r1 + 1) / 2 - off;
        HIP_CHECK(hipMemcpyAsync(stage.p, lt + off, sizeof(float) * cnt, hipMemcpyHostToDevice, s));
        launch_grm_tiles_from_lt(s, stage.as<float>(), r0, r1, c->d_grm.as<float>());
        r0 = r1;
    }
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<double> gd((size_t)n);
    for (int64_t i = 0; i < n; ++i) gd[i] = (double)lt[(size_t)i * (i + 1) / 2 + i];
    c->d_gdiag.ensure(sizeof(double) * n);
    HIP_CHECK(hipMemcpy(c->d_gdiag.p, gd.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    c->gdiag_ok = true;
    c->gen = 2;
    c->gen_n = (int32_t)n;
    c->grm8_ok = false;
    API_END(c)
}

int globe_grm_from_geno(globe_ctx* c) {
    API_BEGIN(c)
    require(c->gen == 1, GLOBE_E_STATE, "genotypes must be loaded to build the GRM");
    const int32_t n = c->n;
    hipStream_t s = c->stream;
    constexpr int kb = 64;  // columns of G per contraction = one tile column
    c->d_grm.ensure(grm_tile_bytes(n));
    HIP_CHECK(hipMemsetAsync(c->d_grm.p, 0, grm_tile_bytes(n), s));
    DevBuf GZ;
    GZ.ensure(sizeof(double) * (size_t)n * kb);
    // G e_j, rows i >= j (W^T e_j is individual j's standardised row; W u runs on 4
    // int8 digits, about 1e-10 relative), rounded to FP32
    for (int64_t j0 = 0; j0 < n; j0 += kb) {
        const int k = (int)std::min<int64_t>(kb, n - j0);
        // rows from j0 down to a multiple of 512 (the packed layouts' row padding); the
        // contraction reuses the block shape tuned for all rows on the shorter ranges
        geno_unit_columns(c, j0, k, GZ.as<double>(), j0 / kRowPad * kRowPad);
        launch_grm_scatter(s, GZ.as<double>(), n, n, j0, k, c->d_grm.as<float>());
    }
    HIP_CHECK(hipStreamSynchronize(s));
    // the context now works from the GRM, as after globe_load_grm
    c->gen = 2;
    c->gen_n = n;
    c->grm8_ok = false;
    c->d_snp.release();
    c->d_ind.release();
    c->m = 0;
    c->d_gdiag.ensure(sizeof(double) * (size_t)n);
    launch_grm_diag(s, c->d_grm.as<float>(), n, c->d_gdiag.as<double>());
    HIP_CHECK(hipStreamSynchronize(s));
    c->gdiag_ok = true;
    API_END(c)
}

int globe_get_grm(globe_ctx* c, float* lt) {
    API_BEGIN(c)
    require(c->gen == 2, GLOBE_E_STATE, "no GRM in the context");
    if (!lt) return GLOBE_OK;
    const int64_t n = c->n;
    hipStream_t s = c->stream;
    DevBuf stage;
    stage.ensure(sizeof(float) * (size_t)std::min<int64_t>(kGrmStageFloats, n * (n + 1) / 2 + n));
    for (int64_t r0 = 0; r0 < n;) {
        const int64_t r1 = grm_chunk_end(r0, n, kGrmStageFloats);
        const int64_t off = r0 * (r0 + 1) / 2, cnt = r1 * (r1 + 1) / 2 - off;
        launch_grm_lt_from_tiles(s, c->d_grm.as<float>(), r0, r1, stage.as<float>());
        HIP_CHECK(hipMemcpyAsync(lt + off, stage.p, sizeof(float) * cnt, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        r0 = r1;
    }
    API_END(c)
}

int globe_get_grm_rows(globe_ctx* c, int64_t r0, int64_t r1, float* rows) {
    API_BEGIN(c)
    require(c->gen == 2, GLOBE_E_STATE, "no GRM in the context");
    require(r0 >= 0 && r0 <= r1 && r1 <= c->n, GLOBE_E_INVALID, "rows must satisfy 0 <= r0 <= r1 <= n");
    require(rows != nullptr || r0 == r1, GLOBE_E_INVALID, "null output");
    hipStream_t s = c->stream;
    DevBuf stage;
    const int64_t base = r0 * (r0 + 1) / 2;
    stage.ensure(sizeof(float) * (size_t)std::min<int64_t>(kGrmStageFloats, r1 * (r1 + 1) / 2 - base + c->n));
    for (int64_t a = r0; a < r1;) {
        const int64_t b = std::min(r1, grm_chunk_end(a, c->n, kGrmStageFloats));
        const int64_t off = a * (a + 1) / 2, cnt = b * (b + 1) / 2 - off;
        launch_grm_lt_from_tiles(s, c->d_grm.as<float>(), a, b, stage.as<float>());
        HIP_CHECK(hipMemcpyAsync(rows + (off - base), stage.p, sizeof(float) * cnt, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        a = b;
    }
    API_END(c)
}

int globe_genetic_diag(globe_ctx* c, double* out) {
    API_BEGIN(c)
    ensure_gdiag(c);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    if (out) HIP_CHECK(hipMemcpy(out, c->d_gdiag.p, sizeof(double) * c->n, hipMemcpyDeviceToHost));
    API_END(c)
}

int globe_set_varcomp(globe_ctx* c, const double theta[5]) {
    API_BEGIN(c)
    require(theta != nullptr, GLOBE_E_INVALID, "null theta");
    for (int q = 0; q < 5; ++q)
        require(std::isfinite(theta[q]), GLOBE_E_INVALID, "non-finite variance component");
    require(theta[0] > 0.0, GLOBE_E_INVALID, "sig2e must be positive");
    for (int q = 0; q < 5; ++q) c->theta[q] = theta[q];
    c->has_theta = true;
    c->theta_k = -1;
    ensure_theta(c, kMaxCols);
    API_END(c)
}

int globe_vx_dev(globe_ctx* c, const double* X, int64_t ldx, int k, double* Y, int64_t ldy,
                 int mode, int precision) {
    API_BEGIN(c)
    require(mode == GLOBE_VX_FULL || mode == GLOBE_VX_GRM_ONLY, GLOBE_E_INVALID, "unknown mode");
    require(precision >= GLOBE_PREC_FP64 && precision <= GLOBE_PREC_FP64_VALU, GLOBE_E_INVALID,
            "unknown precision");
    const int64_t rows_out = mode == GLOBE_VX_FULL ? c->nrec : c->n;
    require(ldx >= c->nrec && ldy >= rows_out, GLOBE_E_INVALID, "leading dimension too small");
    op_vx(c, X, ldx, k, Y, ldy, mode, precision, nullptr, nullptr);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    API_END(c)
}

int globe_vx(globe_ctx* c, const double* X, int64_t ldx, int k, double* Y, int64_t ldy, int mode,
             int precision) {
    API_BEGIN(c)
    require(X && Y && k >= 1 && k <= kMaxCols, GLOBE_E_INVALID, "globe_vx: bad arguments");
    require(c->has_cohort, GLOBE_E_STATE, "cohort not set");
    const int64_t rows_out = mode == GLOBE_VX_FULL ? c->nrec : c->n;
    require(ldx >= c->nrec && ldy >= rows_out, GLOBE_E_INVALID, "leading dimension too small");
    c->s_in.ensure(sizeof(double) * c->nrec * k);
    c->s_out.ensure(sizeof(double) * rows_out * k);
    HIP_CHECK(hipMemcpy2D(c->s_in.p, sizeof(double) * c->nrec, X, sizeof(double) * ldx,
                          sizeof(double) * c->nrec, k, hipMemcpyHostToDevice));
    int rc = globe_vx_dev(c, c->s_in.as<double>(), c->nrec, k, c->s_out.as<double>(), rows_out, mode,
                          precision);
    if (rc != GLOBE_OK) return rc;
    HIP_CHECK(hipMemcpy2D(Y, sizeof(double) * ldy, c->s_out.p, sizeof(double) * rows_out,
                          sizeof(double) * rows_out, k, hipMemcpyDeviceToHost));
    API_END(c)
}

int globe_precond_apply(globe_ctx* c, const double* R, int64_t ldr, int k, double* Z, int64_t ldz) {
    API_BEGIN(c)
    require(R && Z && k >= 1 && k <= kMaxCols && ldr >= c->nrec && ldz >= c->nrec, GLOBE_E_INVALID,
            "globe_precond_apply: bad arguments");
    require(c->has_cohort && c->gen != 0, GLOBE_E_STATE, "cohort and genetics required");
    c->s_in.ensure(sizeof(double) * c->nrec * k);
    c->s_out.ensure(sizeof(double) * c->nrec * k);
    HIP_CHECK(hipMemcpy2D(c->s_in.p, sizeof(double) * c->nrec, R, sizeof(double) * ldr,
                          sizeof(double) * c->nrec, k, hipMemcpyHostToDevice));
    precond(c, c->s_in.as<double>(), c->nrec, k, c->s_out.as<double>(), c->nrec, nullptr, nullptr, nullptr);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    HIP_CHECK(hipMemcpy2D(Z, sizeof(double) * ldz, c->s_out.p, sizeof(double) * c->nrec,
                          sizeof(double) * c->nrec, k, hipMemcpyDeviceToHost));
    API_END(c)
}

int globe_pcg_dev(globe_ctx* c, const double* B, int64_t ldb, int k, double* X, int64_t ldx,
                  const globe_pcg_opts* opts, globe_pcg_stats* stats, double* col_ratio,
                  int32_t* col_iters) {
    API_BEGIN(c)
    require(B && X && opts && ldb >= c->nrec && ldx >= c->nrec, GLOBE_E_INVALID,
            "globe_pcg: bad arguments");
    run_pcg(c, B, ldb, k, X, ldx, *opts, stats, col_ratio, col_iters);
    API_END(c)
}

int globe_pcg(globe_ctx* c, const double* B, int64_t ldb, int k, double* X, int64_t ldx,
              const globe_pcg_opts* opts, globe_pcg_stats* stats, double* col_ratio,
              int32_t* col_iters) {
    API_BEGIN(c)
    require(B && X && opts && k >= 1 && k <= kMaxCols && ldb >= c->nrec && ldx >= c->nrec,
            GLOBE_E_INVALID, "globe_pcg: bad arguments");
    require(c->has_cohort, GLOBE_E_STATE, "cohort not set");
    const int64_t nrec = c->nrec;
    c->s_in.ensure(sizeof(double) * nrec * k);
    c->s_out.ensure(sizeof(double) * nrec * k);
    HIP_CHECK(hipMemcpy2D(c->s_in.p, sizeof(double) * nrec, B, sizeof(double) * ldb,
                          sizeof(double) * nrec, k, hipMemcpyHostToDevice));
    if (opts->x0_mode == GLOBE_X0_GIVEN)
        HIP_CHECK(hipMemcpy2D(c->s_out.p, sizeof(double) * nrec, X, sizeof(double) * ldx,
                              sizeof(double) * nrec, k, hipMemcpyHostToDevice));
    run_pcg(c, c->s_in.as<double>(), nrec, k, c->s_out.as<double>(), nrec, *opts, stats, col_ratio,
            col_iters);
    HIP_CHECK(hipMemcpy2D(X, sizeof(double) * ldx, c->s_out.p, sizeof(double) * nrec,
                          sizeof(double) * nrec, k, hipMemcpyDeviceToHost));
    API_END(c)
}

}  // extern "C"

// ------------------------------------------------------------------------ estimators
namespace globe {

// fitMME's closing steps (PCGsolver.cpp:262-290, varCompEst.cpp:636-665) from a solved
// block in c->e_sol (ld nrec): columns 0..p hold V^-1 X, column p+1 V^-1 y. Forms
// X^T V^-1 [X | y], beta = (X^T V^-1 X)^-1 X^T V^-1 y and V^-1 r = V^-1 y - V^-1 X beta
// into vinv_r_dev (which may be column p+1 itself).
static void fit_finish(globe_ctx* c, Fit& f, double* vinv_r_dev) {
    const int64_t nrec = c->nrec;
    const int q = c->p + 1, k = c->p + 2;
    hipStream_t s = c->stream;
    double* sol = c->e_sol.as<double>();
    c->e_small.ensure(sizeof(double) * kMaxCols * kMaxCols);
    c->w_part.ensure(sizeof(double) * kMaxParts * kMaxCols * kMaxCols);
    launch_gram(s, nrec, c->d_X.as<double>(), nrec, q, sol, nrec, k, c->w_part.as<double>(),
                c->e_small.as<double>());
    std::vector<double> g((size_t)q * k);
    HIP_CHECK(hipMemcpyAsync(g.data(), c->e_small.p, sizeof(double) * q * k, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<double> A((size_t)q * q), Ai;
    for (int a = 0; a < q; ++a)
        for (int b = 0; b < q; ++b) A[a * q + b] = g[a * k + b];
    require(invert(q, A, Ai), GLOBE_E_NUMERIC, "X^T V^-1 X is singular");
    f.beta.assign(q, 0.0);
    for (int a = 0; a < q; ++a)
        for (int b = 0; b < q; ++b) f.beta[a] += Ai[a * q + b] * g[b * k + q];
    f.xtvx_inv = Ai;
    double* vinv_y = sol + (size_t)q * nrec;
    if (vinv_r_dev != vinv_y) launch_copy_cols(s, nrec, vinv_y, nrec, vinv_r_dev, nrec, 1);
    c->e_coef.ensure(sizeof(double) * kMaxCols * kMaxCols);
    HIP_CHECK(hipMemcpyAsync(c->e_coef.p, f.beta.data(), sizeof(double) * q, hipMemcpyHostToDevice, s));
    launch_axpy_cols(s, nrec, vinv_r_dev, nrec, sol, nrec, q, c->e_coef.as<double>(), 1);
    HIP_CHECK(hipStreamSynchronize(s));
}

// fitMME (PCGsolver.cpp:262-290, varCompEst.cpp:636-665): solve V [X | y] into c->e_sol
// (nrec x (p+2)), then beta and V^-1 r into vinv_r_dev.
Fit fit_gls(globe_ctx* c, const double* y_dev, const globe_pcg_opts& o, double* vinv_r_dev) {
    const int64_t nrec = c->nrec;
    const int q = c->p + 1, k = c->p + 2;
    require(k <= kMaxCols, GLOBE_E_INVALID, "too many covariates");
    hipStream_t s = c->stream;
    c->e_xy.ensure(sizeof(double) * nrec * k);
    c->e_sol.ensure(sizeof(double) * nrec * k);
    double* xy = c->e_xy.as<double>();
    double* sol = c->e_sol.as<double>();
    launch_copy_cols(s, nrec, c->d_X.as<double>(), nrec, xy, nrec, q);
    launch_copy_cols(s, nrec, y_dev, nrec, xy + (size_t)q * nrec, nrec, 1);
    Fit f;
    run_pcg(c, xy, nrec, k, sol, nrec, o, &f.stats, nullptr, nullptr);
    fit_finish(c, f, vinv_r_dev);
    return f;
}

// P v = V^-1 v - V^-1 X (X^T V^-1 X)^-1 X^T V^-1 v for ncols columns in place
// (varCompEst.cpp:960, :1028-1035); sv holds V^-1 v, V^-1 X is e_sol[:, :p+1].
void apply_P(globe_ctx* c, const Fit& f, double* sv, int ncols) {
    const int64_t nrec = c->nrec;
    const int q = c->p + 1;
    hipStream_t s = c->stream;
    c->e_small.ensure(sizeof(double) * kMaxCols * kMaxCols);
    c->w_part.ensure(sizeof(double) * kMaxParts * kMaxCols * kMaxCols);
    launch_gram(s, nrec, c->d_X.as<double>(), nrec, q, sv, nrec, ncols, c->w_part.as<double>(),
                c->e_small.as<double>());
    std::vector<double> xv((size_t)q * ncols), coef((size_t)q * ncols, 0.0);
    HIP_CHECK(hipMemcpyAsync(xv.data(), c->e_small.p, sizeof(double) * q * ncols, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (int a = 0; a < q; ++a)
        for (int j = 0; j < ncols; ++j)
            for (int b = 0; b < q; ++b) coef[a * ncols + j] += f.xtvx_inv[a * q + b] * xv[b * ncols + j];
    c->e_coef.ensure(sizeof(double) * kMaxCols * kMaxCols);
    HIP_CHECK(hipMemcpyAsync(c->e_coef.p, coef.data(), sizeof(double) * q * ncols, hipMemcpyHostToDevice, s));
    launch_axpy_cols(s, nrec, sv, nrec, c->e_sol.as<double>(), nrec, q, c->e_coef.as<double>(), ncols);
    HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace globe

extern "C" {

int globe_gls_blup(globe_ctx* c, const double* y, const globe_pcg_opts* opts, double* beta,
                   double* vinv_r, double* uhat, double* bhat, globe_pcg_stats* stats) {
    API_BEGIN(c)
    require(y && opts, GLOBE_E_INVALID, "globe_gls_blup: null argument");
    require(c->has_cohort && c->gen != 0, GLOBE_E_STATE, "cohort and genetics required");
    const int64_t nrec = c->nrec;
    hipStream_t s = c->stream;
    c->e_probe.ensure(sizeof(double) * nrec);
    c->e_pv.ensure(sizeof(double) * nrec);
    HIP_CHECK(hipMemcpy(c->e_probe.p, y, sizeof(double) * nrec, hipMemcpyHostToDevice));
    Fit f = fit_gls(c, c->e_probe.as<double>(), *opts, c->e_pv.as<double>());
    if (beta) std::copy(f.beta.begin(), f.beta.end(), beta);
    if (stats) *stats = f.stats;
    if (vinv_r) HIP_CHECK(hipMemcpy(vinv_r, c->e_pv.p, sizeof(double) * nrec, hipMemcpyDeviceToHost));
    if (uhat) {
        // compute_uhat (PCGsolver.cpp:1447-1473): sig2g G J^T V^-1 r
        c->e_us.ensure(sizeof(double) * c->n);
        op_vx(c, c->e_pv.as<double>(), nrec, 1, c->e_us.as<double>(), c->n, GLOBE_VX_GRM_ONLY,
              opts->precision, nullptr, nullptr);
        HIP_CHECK(hipStreamSynchronize(s));
        HIP_CHECK(hipMemcpy(uhat, c->e_us.p, sizeof(double) * c->n, hipMemcpyDeviceToHost));
    }
    if (bhat) {
        c->e_fs.ensure(sizeof(double) * 2 * c->n);
        launch_bhat(s, c->n, c->d_blk_off.as<int64_t>(), c->d_blk_ind.as<int32_t>(),
                    c->d_times.as<double>(), c->e_pv.as<double>(), c->d_theta.as<double>(),
                    c->e_fs.as<double>());
        HIP_CHECK(hipStreamSynchronize(s));
        HIP_CHECK(hipMemcpy(bhat, c->e_fs.p, sizeof(double) * 2 * c->n, hipMemcpyDeviceToHost));
    }
    API_END(c)
}

int globe_aireml_eval(globe_ctx* c, const double* y, const double* probes, const double* weights,
                      const double* probe_x0, int nmc, const globe_pcg_opts* opts, double* grad,
                      double* ai, double* first_terms, double* trace) {
    if (!c) return GLOBE_E_INVALID;
    if (!y || !grad || !ai) {
        c->err = "globe_aireml_eval: null argument";
        return GLOBE_E_INVALID;
    }
    return globe::aireml_eval_impl(c, y, probes, weights, probe_x0, nmc, opts, true, grad, ai,
                                   first_terms, trace, 5, nullptr);
}

int globe_aireml_eval_reduced(globe_ctx* c, const double* y, const double* probes, const double* weights,
                              const double* probe_x0, int nmc, const globe_pcg_opts* opts, double* grad,
                              double* ai, double* first_terms, double* trace) {
    if (!c) return GLOBE_E_INVALID;
    if (!y || !grad || !ai) {
        c->err = "globe_aireml_eval_reduced: null argument";
        return GLOBE_E_INVALID;
    }
    return globe::aireml_eval_impl(c, y, probes, weights, probe_x0, nmc, opts, true, grad, ai,
                                   first_terms, trace, 2, nullptr);
}

int globe_aireml_probes(globe_ctx* c, int kind, int nmc, uint64_t seed, double* probes, double* weights) {
    API_BEGIN(c)
    require(c->has_cohort && c->has_theta, GLOBE_E_STATE, "cohort and variance components required");
    require(kind >= 0 && kind <= 2, GLOBE_E_INVALID, "unknown probe kind");
    require(nmc >= 1 && nmc <= kMaxCols, GLOBE_E_INVALID, "nmc must be in [1, 64]");
    require(kind != 2 || nmc == c->nrec, GLOBE_E_INVALID, "exact probes need nmc equal to the record count");
    const size_t bytes = sizeof(double) * (size_t)c->nrec * nmc;
    DevBuf Y, Wt;
    Y.ensure(bytes);
    Wt.ensure(bytes);
    make_probes_dev(c, kind, nmc, seed, Y.as<double>(), Wt.as<double>());
    if (probes) HIP_CHECK(hipMemcpyAsync(probes, Y.p, bytes, hipMemcpyDeviceToHost, c->stream));
    if (weights) HIP_CHECK(hipMemcpyAsync(weights, Wt.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    API_END(c)
}

}  // extern "C"

namespace globe {

// computeDerivative_first + estimateTrace_PdV + computeDerivative_AImat
// (varCompEst.cpp:980-1041; reduced model varCompEst_reducedModel_noTime.cpp:662-705).
// y == nullptr reuses the fit, f_theta and X^T V^-1 X of the previous call at the same
// variance components (the AI matrix of an accepted step, varCompEst.cpp:1321-1330,
// needs no new fit). nvc = 5 (full model) or 2 (reduced model: sig2b0 = sigb01 =
// sig2b1 = 0 and one record per individual, so V = sig2g A + sig2e I and f_theta has
// the columns V^-1 r and A V^-1 r, compute_f_theta_i :524-548); grad, first_terms and
// trace hold nvc values, ai nvc x nvc row-major.
// y (nrec) into dst on the device: from the copy the running AI-REML fit uploaded once,
// else from the host.
static void upload_y(globe_ctx* c, const double* y, double* dst) {
    if (y && y == c->yfit_host && c->e_yfit.p) {
        launch_copy_cols(c->stream, c->nrec, c->e_yfit.as<double>(), c->nrec, dst, c->nrec, 1);
        return;
    }
    HIP_CHECK(hipMemcpy(dst, y, sizeof(double) * c->nrec, hipMemcpyHostToDevice));
}

int aireml_eval_impl(globe_ctx* c, const double* y, const double* probes, const double* weights,
                     const double* probe_x0, int nmc, const globe_pcg_opts* opts, bool want_ai,
                     double* grad, double* ai, double* first_terms, double* trace, int nvc,
                     const ProbeGen* gen) {
    API_BEGIN(c)
    require(opts != nullptr, GLOBE_E_INVALID, "globe_aireml_eval: null options");
    require(nvc == 5 || nvc == 2, GLOBE_E_INVALID, "2 or 5 variance components");
    require(!gen || (gen->kind >= 0 && gen->kind <= 2), GLOBE_E_INVALID, "unknown probe kind");
    require(!gen || gen->kind != 2 || nmc == c->nrec, GLOBE_E_INVALID,
            "exact probes need numMC equal to the record count");
    if (nvc == 2) {
        require(c->has_cohort && c->nrec == c->n, GLOBE_E_INVALID,
                "the reduced model takes exactly one record per individual");
        require(c->theta[2] == 0.0 && c->theta[3] == 0.0 && c->theta[4] == 0.0, GLOBE_E_INVALID,
                "the reduced model has no random intercept or slope (sig2b0 = sigb01 = sig2b1 = 0)");
    }
    require(nmc >= 0 && nmc <= kMaxCols, GLOBE_E_INVALID, "nmc must be in [0, 64]");
    require(nmc == 0 || gen || (probes && weights), GLOBE_E_INVALID, "probes/weights required");
    require(c->has_cohort && c->gen != 0, GLOBE_E_STATE, "cohort and genetics required");
    const int64_t nrec = c->nrec;
    const int32_t n = c->n;
    hipStream_t s = c->stream;
    const double sg = c->theta[1];
    require(sg > 0.0, GLOBE_E_INVALID, "sig2g must be positive for AIREML (f_theta divides by it)");
    const int q = c->p + 1, kf = c->p + 2;
    require(kf <= kMaxCols, GLOBE_E_INVALID, "too many covariates");
    // y and the probes are solved as one block when [X | y | probes] fits in one
    const bool merged = y && nmc > 0 && kf + nmc <= kMaxCols;
    globe_pcg_opts o = *opts;
    o.x0_mode = GLOBE_X0_PRECOND;
    c->e_probe.ensure(sizeof(double) * nrec * std::max(1, nmc));
    c->e_pv.ensure(sizeof(double) * nrec * std::max(5, nmc));
    c->e_w.ensure(sizeof(double) * nrec);
    c->e_us.ensure(sizeof(double) * (size_t)n * (1 + nmc));
    c->e_f.ensure(sizeof(double) * nrec * 5);
    c->e_small.ensure(sizeof(double) * kMaxCols * kMaxCols);
    c->w_part.ensure(sizeof(double) * kMaxParts * kMaxCols * kMaxCols);
    if (nmc > 0) {
        c->e_sv.ensure(sizeof(double) * nrec * nmc);
        c->e_fs.ensure(sizeof(double) * nrec * 5 * nmc);
        c->s_aux.ensure(sizeof(double) * nrec * nmc);
    }
    double* py = c->e_w.as<double>();
    double* fth = c->e_f.as<double>();
    double* us = c->e_us.as<double>();      // G-products: column 0 for V^-1 r, 1.. for the probes
    double* wts = c->s_aux.as<double>();    // trace weights of the probes
    double* sv = c->e_sv.as<double>();      // V^-1 probes, then P probes
    globe_pcg_opts ot = o;                  // the probes' solve
    // The probes into Y and their PCG start into X0 (ld nrec). Device probes: weights =
    // tilde-V^-1 probe (also the start of its solve) for the Gaussian tilde-V probes, the
    // probe itself otherwise; injected probes: as given.
    auto load_probes = [&](double* Y, double* X0) {
        if (gen) {
            make_probes_dev(c, gen->kind, nmc, gen->seed, Y, wts);
            if (gen->kind == 0) {
                launch_copy_cols(s, nrec, wts, nrec, X0, nrec, nmc);
                ot.x0_mode = GLOBE_X0_GIVEN;
            }
        } else {
            HIP_CHECK(hipMemcpy(Y, probes, sizeof(double) * nrec * nmc, hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(wts, weights, sizeof(double) * nrec * nmc, hipMemcpyHostToDevice));
            if (probe_x0) {
                HIP_CHECK(hipMemcpy(X0, probe_x0, sizeof(double) * nrec * nmc, hipMemcpyHostToDevice));
                ot.x0_mode = GLOBE_X0_GIVEN;
            }
        }
    };
    Fit f;
    if (merged) {
        // 1-3. fitMME (varCompEst.cpp:587-685) and the probe solves of estimateTrace_PdV
        //    (:816-977) as one block PCG of [X | y | probes], each column from its own x0
        //    (Q^-1 b, or the probe's given start). Under the all-columns rule the fit's
        //    columns iterate until the probes converge too: never looser than a solve of
        //    their own.
        const int kb = kf + nmc;
        c->e_xy.ensure(sizeof(double) * nrec * kb);
        c->e_sol.ensure(sizeof(double) * nrec * kb);
        double* xy = c->e_xy.as<double>();
        double* sol = c->e_sol.as<double>();
        launch_copy_cols(s, nrec, c->d_X.as<double>(), nrec, xy, nrec, q);
        upload_y(c, y, xy + (size_t)q * nrec);
        sv = sol + (size_t)kf * nrec;
        load_probes(xy + (size_t)kf * nrec, sv);
        precond(c, xy, nrec, ot.x0_mode == GLOBE_X0_GIVEN ? kf : kb, sol, nrec, nullptr, nullptr, nullptr);
        globe_pcg_opts om = o;
        om.x0_mode = GLOBE_X0_GIVEN;
        run_pcg(c, xy, nrec, kb, sol, nrec, om, &f.stats, nullptr, nullptr);
        // V^-1 r replaces V^-1 y in column p+1, beside the probes' P z: uhat and the
        // probes' G-products take one GRM-only product over the 1 + nmc columns
        double* vr = sol + (size_t)q * nrec;
        fit_finish(c, f, vr);
        apply_P(c, f, sv, nmc);
        launch_copy_cols(s, nrec, vr, nrec, py, nrec, 1);
        op_vx(c, vr, nrec, 1 + nmc, us, n, GLOBE_VX_GRM_ONLY, o.precision, nullptr, nullptr);
    } else if (y) {
        // 1. fitMME at theta (varCompEst.cpp:587-685): x0 = Q^-1 b
        upload_y(c, y, c->e_probe.as<double>());
        f = fit_gls(c, c->e_probe.as<double>(), o, py);
        op_vx(c, py, nrec, 1, us, n, GLOBE_VX_GRM_ONLY, o.precision, nullptr, nullptr);
    } else {
        require(c->fit_valid && c->fit_theta == c->theta, GLOBE_E_STATE,
                "no fit at the current variance components to reuse");
        f.xtvx_inv = c->fit_xtvx_inv;
    }
    if (y) {
        // 2. uhat and f_theta (varCompEst.cpp:693-760)
        launch_f_theta(s, n, c->d_blk_off.as<int64_t>(), c->d_blk_ind.as<int32_t>(),
                       c->d_times.as<double>(), py, nrec, us, n, 1, 1.0 / sg, fth, nrec);
        c->fit_xtvx_inv = f.xtvx_inv;
        c->fit_theta = c->theta;
        c->fit_valid = true;
    }
    double first[5] = {0, 0, 0, 0, 0};
    launch_gram(s, nrec, py, nrec, 1, fth, nrec, 5, c->w_part.as<double>(), c->e_small.as<double>());
    HIP_CHECK(hipMemcpyAsync(first, c->e_small.p, sizeof(double) * 5, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    // 3. trace terms (varCompEst.cpp:816-977)
    double tr[5] = {0, 0, 0, 0, 0};
    if (nmc > 0) {
        if (!merged) {
            load_probes(c->e_probe.as<double>(), sv);
            run_pcg(c, c->e_probe.as<double>(), nrec, nmc, sv, nrec, ot, nullptr, nullptr, nullptr);
            apply_P(c, f, sv, nmc);
            op_vx(c, sv, nrec, nmc, us + n, n, GLOBE_VX_GRM_ONLY, o.precision, nullptr, nullptr);
        }
        double* fs = c->e_fs.as<double>();
        launch_f_theta(s, n, c->d_blk_off.as<int64_t>(), c->d_blk_ind.as<int32_t>(),
                       c->d_times.as<double>(), sv, nrec, us + n, n, nmc, 1.0 / sg, fs, nrec);
        launch_gram_grouped(s, nrec, wts, nrec, nmc, fs, nrec, 5, c->w_part.as<double>(),
                            c->e_small.as<double>());
        std::vector<double> t((size_t)5 * nmc);
        HIP_CHECK(hipMemcpyAsync(t.data(), c->e_small.p, sizeof(double) * 5 * nmc, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        for (int smp = 0; smp < nmc; ++smp)
            for (int j = 0; j < 5; ++j) tr[j] += t[smp * 5 + j];
        for (int j = 0; j < 5; ++j) tr[j] /= (double)nmc;
    }
    if (grad)
        for (int j = 0; j < nvc; ++j) grad[j] = 0.5 * (first[j] - tr[j]);
    if (first_terms) std::copy(first, first + nvc, first_terms);
    if (trace) std::copy(tr, tr + nvc, trace);
    if (want_ai && ai) {
        // 4. AI matrix (varCompEst.cpp:1015-1041): PCG(f, x0 = Q^-1 f), P f, AI = f^T P f / 2
        //    over the model's nvc columns of f_theta
        double* pf = c->e_pv.as<double>();
        globe_pcg_opts oa = o;
        oa.x0_mode = GLOBE_X0_PRECOND;
        run_pcg(c, fth, nrec, nvc, pf, nrec, oa, nullptr, nullptr, nullptr);
        apply_P(c, f, pf, nvc);
        launch_gram(s, nrec, fth, nrec, nvc, pf, nrec, nvc, c->w_part.as<double>(), c->e_small.as<double>());
        double aim[25];
        HIP_CHECK(hipMemcpyAsync(aim, c->e_small.p, sizeof(double) * nvc * nvc, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        for (int q = 0; q < nvc * nvc; ++q) ai[q] = 0.5 * aim[q];
    }
    API_END(c)
}

}  // namespace globe

extern "C" {

// ------------------------------------------------------------------ device helpers
int globe_dev_alloc(globe_ctx* c, int64_t bytes, void** ptr) {
    API_BEGIN(c)
    require(ptr && bytes >= 0, GLOBE_E_INVALID, "globe_dev_alloc: bad arguments");
    *ptr = nullptr;
    hipError_t e = hipMalloc(ptr, bytes > 0 ? (size_t)bytes : 256);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(GLOBE_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    API_END(c)
}

int globe_dev_free(globe_ctx* c, void* ptr) {
    API_BEGIN(c)
    HIP_CHECK(hipStreamSynchronize(c->stream));
    if (ptr) HIP_CHECK(hipFree(ptr));
    API_END(c)
}

int globe_memcpy_h2d(globe_ctx* c, void* dst, const void* src, int64_t bytes) {
    API_BEGIN(c)
    require(dst && src && bytes >= 0, GLOBE_E_INVALID, "globe_memcpy_h2d: bad arguments");
    HIP_CHECK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    API_END(c)
}

int globe_memcpy_d2h(globe_ctx* c, void* dst, const void* src, int64_t bytes) {
    API_BEGIN(c)
    require(dst && src && bytes >= 0, GLOBE_E_INVALID, "globe_memcpy_d2h: bad arguments");
    HIP_CHECK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    API_END(c)
}

int globe_device_memory(globe_ctx* c, int64_t* free_bytes, int64_t* total_bytes) {
    API_BEGIN(c)
    size_t f = 0, t = 0;
    HIP_CHECK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    API_END(c)
}

int globe_force_variant(globe_ctx* c, int index) {
    API_BEGIN(c)
    require(index >= -1, GLOBE_E_INVALID, "variant index must be >= -1");
    c->force_variant = index;
    API_END(c)
}

int globe_variant_counts(globe_ctx* c, int32_t counts[2]) {
    API_BEGIN(c)
    require(counts != nullptr, GLOBE_E_INVALID, "null counts");
    counts[0] = c->variant_count[0];
    counts[1] = c->variant_count[1];
    API_END(c)
}

int globe_set_mixed_digits(globe_ctx* c, int digits) {
    API_BEGIN(c)
    require(digits >= 2 && digits <= kLimbsMixed, GLOBE_E_INVALID, "mixed digits must be 2 to 4");
    c->limbs_default = digits;
    c->limbs_mixed = digits;
    API_END(c)
}

int globe_set_graphs(globe_ctx* c, int on) {
    API_BEGIN(c)
    c->graphs_on = on != 0;
    API_END(c)
}

int globe_set_row_gather(globe_ctx* c, int on) {
    API_BEGIN(c)
    c->gather_on = on != 0;
    API_END(c)
}

int globe_profile_enable(globe_ctx* c, int on) {
    API_BEGIN(c)
    HIP_CHECK(hipStreamSynchronize(c->stream));
    for (auto& e : c->prof) {
        if (!e.a) continue;
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    c->prof.clear();
    c->prof_on = on != 0;
    API_END(c)
}

int globe_profile_read(globe_ctx* c, int which, globe_kernel_stats* out) {
    API_BEGIN(c)
    require(out != nullptr, GLOBE_E_INVALID, "null output");
    HIP_CHECK(hipStreamSynchronize(c->stream));
    globe_kernel_stats st{};
    for (const auto& e : c->prof) {
        if (e.pass != which) continue;
        float ms = e.ms;
        if (e.a) HIP_CHECK(hipEventElapsedTime(&ms, e.a, e.b));
        st.launches += 1;
        st.ms_total += ms;
        st.bytes_per_launch = e.bytes;
        st.flops_per_launch += e.flops;
        st.issued_ops_per_launch += e.issued;
    }
    if (st.launches) {
        st.flops_per_launch /= (double)st.launches;
        st.issued_ops_per_launch /= (double)st.launches;
    }
    *out = st;
    API_END(c)
}

int globe_dev_copy(globe_ctx* c, void* dst, const void* src, int64_t bytes) {
    API_BEGIN(c)
    require(dst && src && bytes >= 0 && bytes % 16 == 0, GLOBE_E_INVALID,
            "globe_dev_copy: bytes must be a non-negative multiple of 16");
    launch_copy16(c->stream, dst, src, bytes);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    API_END(c)
}

int globe_synchronize(globe_ctx* c) {
    API_BEGIN(c)
    HIP_CHECK(hipStreamSynchronize(c->stream));
    API_END(c)
}

int globe_geno_Wa(globe_ctx* c, const double* a, double* out) {
    API_BEGIN(c)
    require(a && out, GLOBE_E_INVALID, "globe_geno_Wa: null argument");
    require(c->gen == 1, GLOBE_E_STATE, "no genotypes loaded");
    const int64_t m = c->m;
    const int32_t n = c->n;
    hipStream_t s = c->stream;
    // out_i = sum_s g_is inv_s a_s - sum_s 2 f_s inv_s a_s: scale a on the host, then
    // one pass over the individual-major rows.
    std::vector<double> inv((size_t)m), twof((size_t)m), u((size_t)m);
    HIP_CHECK(hipMemcpy(inv.data(), c->d_inv.p, sizeof(double) * m, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(twof.data(), c->d_twof.p, sizeof(double) * m, hipMemcpyDeviceToHost));
    double csum = 0.0;
    for (int64_t q = 0; q < m; ++q) {
        u[q] = inv[q] * a[q];
        csum += twof[q] * u[q];
    }
    c->w_U.ensure(sizeof(double) * std::max<int64_t>(m, 1));
    c->w_P2.ensure(sizeof(double) * n);
    HIP_CHECK(hipMemcpy(c->w_U.p, u.data(), sizeof(double) * m, hipMemcpyHostToDevice));
    launch_rowdot(s, c->d_ind.as<uint8_t>(), c->pitch_i, n, m, c->w_U.as<double>(), m, 1,
                  c->w_P2.as<double>(), n, nullptr);
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipMemcpy(out, c->w_P2.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < n; ++i) out[i] -= csum;
    API_END(c)
}

}  // extern "C"
