// AI-REML outer loop and trust-region step (host FP64; the derivatives, trace and
// AI matrix come from the device, aireml_eval_impl). Two models share the code:
//   full    -- five components (sig2e, sig2g, sig2b0, sigb01, sig2b1) of the
//              longitudinal model (varCompEst.cpp);
//   reduced -- two components (sig2e, sig2g), V = sig2g A + sig2e I with one record per
//              individual and no time term (varCompEst_reducedModel_noTime.cpp). On the
//              device it is the full operator with sig2b0 = sigb01 = sig2b1 = 0 and zero
//              times, whose per-individual block preconditioner is then the diagonal
//              1 / (sig2g A_ii + sig2e) of computePreconditioner (:326-335).
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "context.h"

using globe::Error;

using globe::kMaxCols;

namespace {

constexpr int KMAX = 5;  // components of the full model; matrices use this stride

// Trust-region subproblem of oneIter_AIREML: minimise q(p) = -(g.p - p^T A p / 2) over the
// box lb <= p <= ub subject to the general constraints c_j(p) <= 0:
//   j = 0: sum_i A_ii p_i - delta (the reference's signed "trust region", varCompEst.cpp
//          :1068-1089, reduced :731-757);
//   j = 1: (full model) (th3+p3)^2 - (th2+p2)(th4+p4) + 1e-3 (covariance, :1121-1145).
// Positivity (:1096-1115, reduced :759-781) is part of the box.
struct StepProblem {
    int nv = KMAX;
    bool cov = true;
    double g[KMAX], A[KMAX * KMAX], th[KMAX], delta;
    double lb[KMAX], ub[KMAX];
    int ncons() const { return cov ? 2 : 1; }
    double cons(int j, const double* p) const {
        if (j == 0) {
            double s = 0.0;
            for (int i = 0; i < nv; ++i) s += A[i * KMAX + i] * p[i];
            return s - delta;
        }
        const double b0 = th[2] + p[2], b01 = th[3] + p[3], b1 = th[4] + p[4];
        return b01 * b01 - b0 * b1 + 1e-3;
    }
    void cons_grad(int j, const double* p, double* gr) const {
        for (int i = 0; i < nv; ++i) gr[i] = 0.0;
        if (j == 0) {
            for (int i = 0; i < nv; ++i) gr[i] = A[i * KMAX + i];
            return;
        }
        const double b0 = th[2] + p[2], b01 = th[3] + p[3], b1 = th[4] + p[4];
        gr[2] = -b1;
        gr[3] = 2.0 * b01;
        gr[4] = -b0;
    }
    bool feasible(const double* p, double tol) const {
        for (int i = 0; i < nv; ++i)
            if (p[i] < lb[i] - tol || p[i] > ub[i] + tol) return false;
        if (cons(0, p) > tol * (1.0 + std::fabs(delta))) return false;
        return !cov || cons(1, p) <= tol;
    }
    double quad(const double* p) const {
        double gp = 0.0, pap = 0.0;
        for (int i = 0; i < nv; ++i) {
            gp += g[i] * p[i];
            for (int j = 0; j < nv; ++j) pap += p[i] * A[i * KMAX + j] * p[j];
        }
        return -(gp - 0.5 * pap);
    }
};

// Augmented Lagrangian value and gradient (box handled by projection).
double al_value(const StepProblem& P, const double* p, const double* lam, double mu, double* grad) {
    const int nv = P.nv;
    double v = P.quad(p);
    for (int i = 0; i < nv; ++i) {
        double ap = 0.0;
        for (int j = 0; j < nv; ++j) ap += P.A[i * KMAX + j] * p[j];
        grad[i] = -(P.g[i] - ap);
    }
    for (int j = 0; j < P.ncons(); ++j) {
        const double t = P.cons(j, p) + lam[j] / mu;
        if (t > 0.0) {
            v += 0.5 * mu * t * t - lam[j] * lam[j] / (2.0 * mu);
            double cg[KMAX];
            P.cons_grad(j, p, cg);
            for (int i = 0; i < nv; ++i) grad[i] += mu * t * cg[i];
        } else {
            v -= lam[j] * lam[j] / (2.0 * mu);
        }
    }
    return v;
}

void project(const StepProblem& P, double* p) {
    for (int i = 0; i < P.nv; ++i) p[i] = std::fmin(P.ub[i], std::fmax(P.lb[i], p[i]));
}

// Hessian of the augmented Lagrangian at p (constraints with c_j + lam_j / mu > 0).
void al_hessian(const StepProblem& P, const double* p, const double* lam, double mu, double* H) {
    const int nv = P.nv;
    std::memcpy(H, P.A, sizeof(double) * KMAX * KMAX);
    for (int j = 0; j < P.ncons(); ++j) {
        const double t = P.cons(j, p) + lam[j] / mu;
        if (t <= 0.0) continue;
        double cg[KMAX];
        P.cons_grad(j, p, cg);
        for (int a = 0; a < nv; ++a)
            for (int b = 0; b < nv; ++b) H[a * KMAX + b] += mu * cg[a] * cg[b];
        if (j == 1) {  // second derivatives of (b01^2 - b0 b1)
            H[3 * KMAX + 3] += mu * t * 2.0;
            H[2 * KMAX + 4] -= mu * t;
            H[4 * KMAX + 2] -= mu * t;
        }
    }
}

// Solve H_FF d_F = -g_F on the free set by Cholesky, adding a Levenberg shift until
// H_FF + shift I is positive definite.
bool newton_dir(int nv, const double* H, const double* g, const bool* fr, double* d) {
    int idx[KMAX], m = 0;
    for (int i = 0; i < nv; ++i) {
        d[i] = 0.0;
        if (fr[i]) idx[m++] = i;
    }
    if (m == 0) return false;
    double scale = 0.0;
    for (int a = 0; a < m; ++a) scale = std::fmax(scale, std::fabs(H[idx[a] * KMAX + idx[a]]));
    double shift = 0.0;
    for (int attempt = 0; attempt < 40; ++attempt) {
        double L[KMAX][KMAX] = {};
        bool ok = true;
        for (int a = 0; a < m && ok; ++a)
            for (int b = 0; b <= a; ++b) {
                double s = H[idx[a] * KMAX + idx[b]] + (a == b ? shift : 0.0);
                for (int q = 0; q < b; ++q) s -= L[a][q] * L[b][q];
                if (a == b) {
                    if (s <= 0.0) { ok = false; break; }
                    L[a][a] = std::sqrt(s);
                } else {
                    L[a][b] = s / L[b][b];
                }
            }
        if (ok) {
            double y[KMAX];
            for (int a = 0; a < m; ++a) {
                double s = -g[idx[a]];
                for (int q = 0; q < a; ++q) s -= L[a][q] * y[q];
                y[a] = s / L[a][a];
            }
            for (int a = m - 1; a >= 0; --a) {
                double s = y[a];
                for (int q = a + 1; q < m; ++q) s -= L[q][a] * d[idx[q]];
                d[idx[a]] = s / L[a][a];
            }
            return true;
        }
        shift = shift == 0.0 ? 1e-10 * (1.0 + scale) : shift * 10.0;
    }
    return false;
}

bool solve_al(const StepProblem& P, double* p) {
    const int nv = P.nv;
    double lam[2] = {0.0, 0.0}, mu = 10.0;
    project(P, p);
    double prev_viol = 1e300;
    for (int outer = 0; outer < 60; ++outer) {
        // projected Newton on the box-constrained augmented Lagrangian
        double gr[KMAX];
        double f = al_value(P, p, lam, mu, gr);
        for (int it = 0; it < 500; ++it) {
            bool fr[KMAX];
            double pg = 0.0;
            for (int i = 0; i < nv; ++i) {
                const bool at_lo = p[i] <= P.lb[i] + 1e-14 && gr[i] > 0.0;
                const bool at_hi = p[i] >= P.ub[i] - 1e-14 && gr[i] < 0.0;
                fr[i] = !(at_lo || at_hi);
                if (fr[i]) pg = std::fmax(pg, std::fabs(gr[i]));
            }
            if (pg < 1e-13) break;
            double H[KMAX * KMAX], d[KMAX];
            al_hessian(P, p, lam, mu, H);
            if (!newton_dir(nv, H, gr, fr, d)) break;
            double a = 1.0, pn[KMAX], grn[KMAX], fn = f;
            bool ok = false;
            for (int bt = 0; bt < 60; ++bt) {
                for (int i = 0; i < nv; ++i) pn[i] = p[i] + a * d[i];
                project(P, pn);
                fn = al_value(P, pn, lam, mu, grn);
                double dec = 0.0;
                for (int i = 0; i < nv; ++i) dec += gr[i] * (p[i] - pn[i]);
                if (fn <= f - 1e-4 * std::fabs(dec) || fn < f) {
                    ok = true;
                    break;
                }
                a *= 0.5;
            }
            if (!ok) break;
            double step = 0.0;
            for (int i = 0; i < nv; ++i) step = std::fmax(step, std::fabs(pn[i] - p[i]));
            std::memcpy(p, pn, sizeof(double) * nv);
            std::memcpy(gr, grn, sizeof(double) * nv);
            f = fn;
            if (step < 1e-16) break;
        }
        // stop at a KKT point: feasible and multipliers no longer moving (complementary
        // slackness), otherwise update the multipliers and, if feasibility stalls, mu
        double viol = 0.0, dlam = 0.0, lmax = 0.0;
        for (int j = 0; j < P.ncons(); ++j) {
            const double cj = P.cons(j, p);
            viol = std::fmax(viol, cj / (j == 0 ? 1.0 + std::fabs(P.delta) : 1.0));
            const double nl = std::fmax(0.0, lam[j] + mu * cj);
            dlam = std::fmax(dlam, std::fabs(nl - lam[j]));
            lam[j] = nl;
            lmax = std::fmax(lmax, nl);
        }
        if (viol <= 1e-12 && dlam <= 1e-10 * (1.0 + lmax)) return true;
        if (viol > 0.25 * prev_viol) mu *= 10.0;
        prev_viol = viol;
        if (mu > 1e14) break;
    }
    return P.feasible(p, 1e-9);
}

// Gauss-Jordan inverse of an n x n matrix (n <= 5) stored with row stride `ld`.
bool invert_n(int n, const double* a, int ld, double* out, int ldo) {
    double w[KMAX][2 * KMAX];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            w[i][j] = a[i * ld + j];
            w[i][n + j] = i == j;
        }
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(w[r][c]) > std::fabs(w[piv][c])) piv = r;
        if (w[piv][c] == 0.0) return false;
        for (int j = 0; j < 2 * n; ++j) std::swap(w[c][j], w[piv][j]);
        const double d = w[c][c];
        for (int j = 0; j < 2 * n; ++j) w[c][j] /= d;
        for (int r = 0; r < n; ++r)
            if (r != c) {
                const double f = w[r][c];
                for (int j = 0; j < 2 * n; ++j) w[r][j] -= f * w[c][j];
            }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) out[i * ldo + j] = w[i][n + j];
    return true;
}

// The step of oneIter_AIREML for nv components (5: full, 2: reduced). ai is nv x nv
// row-major.
int step_impl(int nv, const double* grad, const double* ai, const double* theta, double delta,
              double* step) {
    StepProblem P;
    P.nv = nv;
    P.cov = nv == KMAX;
    for (int i = 0; i < nv; ++i) {
        P.g[i] = grad[i];
        P.th[i] = theta[i];
        for (int j = 0; j < nv; ++j) P.A[i * KMAX + j] = ai[i * nv + j];
    }
    P.delta = delta;
    for (int i = 0; i < nv; ++i) {
        P.ub[i] = 10.0;
        // positivity for every component but the covariance sigb01 of the full model
        P.lb[i] = (P.cov && i == 3) ? -10.0 : std::fmax(-10.0, 1e-5 - theta[i]);
    }
    double Ai[KMAX * KMAX], p[KMAX];
    if (invert_n(nv, P.A, KMAX, Ai, KMAX)) {
        for (int i = 0; i < nv; ++i) {
            p[i] = 0.0;
            for (int j = 0; j < nv; ++j) p[i] += Ai[i * KMAX + j] * grad[j];
        }
        bool spd = true;  // the Newton point minimises q only when AI is positive definite
        for (int i = 0; i < nv; ++i) spd &= P.A[i * KMAX + i] > 0.0;
        if (spd && P.feasible(p, 0.0)) {
            std::memcpy(step, p, sizeof(double) * nv);
            return GLOBE_OK;
        }
    } else {
        for (int i = 0; i < nv; ++i) p[i] = 0.0;
    }
    // The covariance constraint is not convex, so local minimisers can differ. NLopt
    // starts from the zero step (:1163); the zero step and three points along the
    // (projected) Newton direction are tried here and the best feasible point is kept.
    double starts[4][KMAX];
    for (int i = 0; i < nv; ++i) {
        starts[0][i] = 0.0;
        starts[1][i] = std::fmin(P.ub[i], std::fmax(P.lb[i], p[i]));
        starts[2][i] = 0.5 * starts[1][i];
        starts[3][i] = 0.1 * starts[1][i];
    }
    bool found = false;
    double best[KMAX], fbest = 0.0;
    for (int s = 0; s < 4; ++s) {
        double q[KMAX];
        std::memcpy(q, starts[s], sizeof(double) * nv);
        for (int i = 0; i < nv; ++i)
            if (!std::isfinite(q[i])) q[i] = 0.0;
        if (!solve_al(P, q)) continue;
        const double f = P.quad(q);
        if (!found || f < fbest) {
            found = true;
            fbest = f;
            std::memcpy(best, q, sizeof(double) * nv);
        }
    }
    if (!found) return GLOBE_E_NUMERIC;
    std::memcpy(step, best, sizeof(double) * nv);
    return GLOBE_OK;
}

struct Logger {
    globe_log_fn fn;
    void* user;
    void operator()(const char* fmt, ...) const __attribute__((format(printf, 2, 3)));
};

void Logger::operator()(const char* fmt, ...) const {
    if (!fn) return;
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    fn(buf, user);
}

// splitmix64 finaliser (Steele, Lea and Flood, OOPSLA 2014): decorrelates nearby seeds
uint64_t mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

// Philox key of evaluation `eval_id` of a fit seeded with `seed`.
extern "C" uint64_t globe_probe_seed(uint64_t seed, uint64_t eval_id) { return mix(seed + 0x9E37ull * (eval_id + 1)); }

namespace {

uint64_t probe_seed(uint64_t seed, uint64_t eval_id) { return globe_probe_seed(seed, eval_id); }

}  // namespace

extern "C" int globe_aireml_step(const double grad[5], const double ai[25], const double theta[5],
                                 double delta, double step[5]) {
    if (!grad || !ai || !theta || !step) return GLOBE_E_INVALID;
    return step_impl(KMAX, grad, ai, theta, delta, step);
}

extern "C" int globe_aireml_step_reduced(const double grad[2], const double ai[4],
                                         const double theta[2], double delta, double step[2]) {
    if (!grad || !ai || !theta || !step) return GLOBE_E_INVALID;
    return step_impl(2, grad, ai, theta, delta, step);
}

namespace {

struct Eval {
    double grad[KMAX], ai[KMAX * KMAX];  // grad (nv), AI nv x nv row-major
};

// theta of nv components -> the five of the device operator
void full_theta(int nv, const double* th, double* th5) {
    for (int j = 0; j < KMAX; ++j) th5[j] = j < nv ? th[j] : 0.0;
}

// One gradient (and AI) evaluation at th with fresh probes: generated on the device from
// a Philox key that mixes the fit's seed with the evaluation's number (k_probe.hip;
// globe_aireml_probes returns the same probes). The reference draws new Gaussian tilde-V
// probes for every evaluation too, from std::random_device (varCompEst.cpp:763-797).
void evaluate(globe_ctx* c, const double* y, int nv, const double* th, const globe_aireml_opts& o,
              uint64_t eval_id, bool want_ai, Eval& e) {
    double th5[KMAX];
    full_theta(nv, th, th5);
    int rc = globe_set_varcomp(c, th5);
    if (rc != GLOBE_OK) throw Error(rc, c->err);
    const globe::ProbeGen gen{o.probes, probe_seed(o.seed, eval_id)};
    globe_pcg_opts po = o.pcg;
    rc = globe::aireml_eval_impl(c, y, nullptr, nullptr, nullptr, o.num_mc, &po, want_ai, e.grad, e.ai,
                                 nullptr, nullptr, nv, &gen);
    if (rc != GLOBE_OK) throw Error(rc, c->err);
}

double norm_n(int nv, const double* v) {
    double s = 0;
    for (int i = 0; i < nv; ++i) s += v[i] * v[i];
    return std::sqrt(s);
}

const char* const kNames[KMAX] = {"sig2e", "sig2g", "sig2b0", "sigb01", "sig2b1"};

std::string theta_line(const char* head, int nv, const double* th) {
    std::string s = head;
    char buf[64];
    for (int j = 0; j < nv; ++j) {
        std::snprintf(buf, sizeof(buf), "%s%s = %g", j ? ", " : "  ", kNames[j], th[j]);
        s += buf;
    }
    return s;
}

// runAIREML's main loop + oneIter_AIREML (varCompEst.cpp:1148-1491; reduced model
// varCompEst_reducedModel_noTime.cpp:784-1099) for nv components.
int fit_impl(globe_ctx* c, const double* y, const globe_aireml_opts* opts, int nv, double* theta,
             double* se_out, double* ai_out, int32_t* iters_out, double* sec_per_iter,
             globe_log_fn logfn, void* user) {
    const bool full = nv == KMAX;
    try {
        const globe_aireml_opts& o = *opts;
        Logger log{logfn, user};
        globe::require(o.num_mc >= 1 && o.num_mc <= kMaxCols, GLOBE_E_INVALID,
                       "numMC must be in [1, 64]");
        globe::require(o.probes >= 0 && o.probes <= 2, GLOBE_E_INVALID, "unknown probe kind");
        if (!full)
            globe::require(c->has_cohort && c->nrec == c->n, GLOBE_E_INVALID,
                           "the reduced model takes exactly one record per individual");
        // y stays fixed for the whole fit: one upload, device copies in every evaluation
        globe::require(c->has_cohort, GLOBE_E_STATE, "no cohort");
        c->e_yfit.ensure(sizeof(double) * (size_t)c->nrec);
        HIP_CHECK(hipMemcpy(c->e_yfit.p, y, sizeof(double) * (size_t)c->nrec, hipMemcpyHostToDevice));
        c->yfit_host = y;
        struct YfitGuard {
            globe_ctx* c;
            ~YfitGuard() { c->yfit_host = nullptr; }
        } yfit_guard{c};
        double th[KMAX];
        std::memcpy(th, theta, sizeof(double) * nv);
        Eval cur{};
        double Delta = 1e20;
        bool dangerous = false, grad_check = true;
        int num_rej = 0;
        double gnt = 0.0;
        uint64_t eval_id = 0;
        std::mt19937_64 restart_rng(mix(o.seed ^ 0x5EEDull));
        double total_secs = 0.0;
        int it_done = 0;
        log("Running AI-REML...");
        log("%s", theta_line("Initial variance components:", nv, th).c_str());
        for (int it = 0; it < o.max_iter; ++it) {
            log("---------- Iteration %d ----------", it + 1);
            const auto t0 = std::chrono::steady_clock::now();
            if (it == 0) evaluate(c, y, nv, th, o, eval_id++, true, cur);
            if (it == 1) {
                gnt = 0.1 * norm_n(nv, cur.grad);
                log("Gradient norm threshold: %g", gnt);
            }
            if (num_rej > 5) {
                // random restart in [0.1, 0.9] (varCompEst.cpp:1431-1451, reduced :1044-1057),
                // seeded here
                log("Too many rejections at current point. Re-initalizing optimization.");
                std::uniform_real_distribution<double> u(0.0, 1.0);
                for (int j = 0; j < nv; ++j) th[j] = 0.1 + 0.8 * u(restart_rng);
                if (full && th[3] > std::sqrt(th[2] * th[4])) th[3] = std::sqrt(th[2] * th[4]) - 1e-3;
                evaluate(c, y, nv, th, o, eval_id++, true, cur);
                num_rej = 0;
            }
            double th_old[KMAX];
            std::memcpy(th_old, th, sizeof(double) * nv);
            // ---- oneIter_AIREML (varCompEst.cpp:1148-1344, reduced :784-964)
            double p[KMAX];
            int rc = step_impl(nv, cur.grad, cur.ai, th, Delta, p);
            if (rc != GLOBE_OK) {
                log("Trust-region step failed. Reducing trust region radius.");
                Delta *= 0.75;
                dangerous = true;
                num_rej += 1;
            } else {
                std::string ps = "Predicted step:";
                for (int j = 0; j < nv; ++j) {
                    char b[32];
                    std::snprintf(b, sizeof(b), " %g", p[j]);
                    ps += b;
                }
                log("%s", ps.c_str());
                double scale = 1.0;
                bool accepted = false, converged_pred = false;
                int red = 0;
                while (!accepted && red <= 10) {
                    double sp[KMAX], tn[KMAX];
                    for (int j = 0; j < nv; ++j) {
                        sp[j] = scale * p[j];
                        tn[j] = th[j] + sp[j];
                    }
                    bool violated = false;
                    for (int j = 0; j < nv; ++j) {
                        if (!(full && j == 3) && tn[j] < 1e-6) {
                            tn[j] = 1e-6;
                            violated = true;
                        }
                        if (tn[j] > 10.0) {
                            tn[j] = 10.0;
                            violated = true;
                        }
                    }
                    if (full && tn[3] > std::sqrt(tn[2] * tn[4])) {
                        tn[3] = std::sqrt(tn[2] * tn[4]) - 1e-6;
                        violated = true;
                    }
                    if (violated) {
                        log("Constraints violated at scale %g. Reducing step size.", scale);
                        scale *= 0.5;
                        red += 1;
                        continue;
                    }
                    Eval nw{};
                    evaluate(c, y, nv, tn, o, eval_id++, false, nw);
                    double dll = 0.0, pred = 0.0, spa = 0.0;
                    for (int a = 0; a < nv; ++a) {
                        dll += sp[a] * 0.5 * (cur.grad[a] + nw.grad[a]);
                        pred += cur.grad[a] * sp[a];
                        for (int b = 0; b < nv; ++b) spa += sp[a] * cur.ai[a * nv + b] * sp[b];
                    }
                    pred -= 0.5 * spa;
                    log("Predicted change in log-likelihood (scaled): %g", pred);
                    log("(Approximate) Actual change in log-likelihood: %g", dll);
                    if (scale == 1.0 && pred < 1e-2) {
                        log("Converged: predicted log-likelihood improvement below threshold.");
                        dangerous = false;
                        converged_pred = true;
                        break;
                    }
                    if (grad_check && norm_n(nv, cur.grad) > gnt &&
                        norm_n(nv, nw.grad) > 2 * norm_n(nv, cur.grad)) {
                        log("Dangerous model deviation detected at scale %g. Reducing step size.", scale);
                        scale *= 0.5;
                        red += 1;
                        continue;
                    }
                    const double rho = dll / pred;
                    log("Rho: %g", rho);
                    double npd = 0.0;
                    for (int j = 0; j < nv; ++j) npd += std::fabs(cur.ai[j * nv + j] * sp[j]);
                    if (rho < 0.25) Delta = 0.25 * npd;
                    else if (rho > 0.99) Delta = std::fmax(Delta, 3.5 * npd);
                    if (rho > 0.25) {
                        std::memcpy(th, tn, sizeof(double) * nv);
                        std::memcpy(cur.grad, nw.grad, sizeof(double) * nv);
                        // AI at the accepted point from the trial's fit and f_theta
                        // (computeDerivative_AImat, :1321-1330)
                        globe_pcg_opts po = o.pcg;
                        const int arc = globe::aireml_eval_impl(c, nullptr, nullptr, nullptr, nullptr, 0, &po,
                                                                true, nullptr, cur.ai, nullptr, nullptr, nv,
                                                                nullptr);
                        if (arc != GLOBE_OK) throw Error(arc, c->err);
                        num_rej = 0;
                        accepted = true;
                    } else {
                        log("Step rejected at scale %g, rho = %g. Reducing step size.", scale, rho);
                        scale *= 0.5;
                        red += 1;
                    }
                }
                if (!accepted && !converged_pred) {
                    log("Failed to accept a step after %d reductions.", red);
                    Delta *= 0.75;
                    dangerous = true;
                    num_rej += 1;
                }
            }
            log("%s", theta_line("Estimated variance components:", nv, th).c_str());
            const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            total_secs += secs;
            it_done = it + 1;
            log("Iteration %d took %g seconds.", it + 1, secs);
            double dn = 0.0, on = 0.0;
            for (int j = 0; j < nv; ++j) {
                dn += (th[j] - th_old[j]) * (th[j] - th_old[j]);
                on += th_old[j] * th_old[j];
            }
            const double delta = std::sqrt(dn) / std::sqrt(on);
            if (delta < o.tol && !dangerous) {
                log("Convergence reached after %d iterations.", it + 1);
                break;
            }
            if (dangerous) dangerous = false;
            log("Convergence check: %g", delta);
            log("Trust region radius: %g", Delta);
            if (Delta < 1e-4) {
                log("Trust region radius too small. Terminating optimization.");
                break;
            }
        }
        // SEs from the AI matrix at the final point (:1484-1491, reduced :1084-1091)
        double Ai[KMAX * KMAX];
        const bool inv_ok = invert_n(nv, cur.ai, nv, Ai, KMAX);
        log("Final variance components +/- SEs:");
        for (int j = 0; j < nv; ++j) {
            const double s = inv_ok ? std::sqrt(Ai[j * KMAX + j]) : NAN;
            if (se_out) se_out[j] = s;
            log("  %6s = %g +/- %g", kNames[j], th[j], s);
        }
        std::memcpy(theta, th, sizeof(double) * nv);
        if (ai_out) std::memcpy(ai_out, cur.ai, sizeof(double) * nv * nv);
        if (iters_out) *iters_out = it_done;
        if (sec_per_iter) *sec_per_iter = it_done ? total_secs / it_done : 0.0;
        return GLOBE_OK;
    } catch (const Error& e) {
        c->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return GLOBE_E_INVALID;
    }
}

}  // namespace

extern "C" int globe_aireml_fit(globe_ctx* c, const double* y, const globe_aireml_opts* opts,
                                double theta[5], double se_out[5], double ai_out[25],
                                int32_t* iters_out, double* sec_per_iter, globe_log_fn logfn,
                                void* user) {
    if (!c) return GLOBE_E_INVALID;
    if (!y || !opts || !theta) {
        c->err = "globe_aireml_fit: null argument";
        return GLOBE_E_INVALID;
    }
    return fit_impl(c, y, opts, KMAX, theta, se_out, ai_out, iters_out, sec_per_iter, logfn, user);
}

extern "C" int globe_aireml_fit_reduced(globe_ctx* c, const double* y, const globe_aireml_opts* opts,
                                        double theta[2], double se_out[2], double ai_out[4],
                                        int32_t* iters_out, double* sec_per_iter, globe_log_fn logfn,
                                        void* user) {
    if (!c) return GLOBE_E_INVALID;
    if (!y || !opts || !theta) {
        c->err = "globe_aireml_fit_reduced: null argument";
        return GLOBE_E_INVALID;
    }
    return fit_impl(c, y, opts, 2, theta, se_out, ai_out, iters_out, sec_per_iter, logfn, user);
}
