// AI-REML outer loop and trust-region step (host FP64; the derivatives, trace and
// AI matrix come from globe_aireml_eval on the device).
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

constexpr int K5 = 5;

double quad(const double* g, const double* A, const double* p) {
    double gp = 0.0, pap = 0.0;
    for (int i = 0; i < K5; ++i) {
        gp += g[i] * p[i];
        for (int j = 0; j < K5; ++j) pap += p[i] * A[i * K5 + j] * p[j];
    }
    return -(gp - 0.5 * pap);
}

struct StepProblem {
    double g[K5], A[K5 * K5], th[K5], delta;
    double lb[K5], ub[K5];
    // general constraints c_j(p) <= 0: j = 0 trust region (signed), j = 1 covariance
    double cons(int j, const double* p) const {
        if (j == 0) {
            double s = 0.0;
            for (int i = 0; i < K5; ++i) s += A[i * K5 + i] * p[i];
            return s - delta;
        }
        const double b0 = th[2] + p[2], b01 = th[3] + p[3], b1 = th[4] + p[4];
        return b01 * b01 - b0 * b1 + 1e-3;
    }
    void cons_grad(int j, const double* p, double* gr) const {
        for (int i = 0; i < K5; ++i) gr[i] = 0.0;
        if (j == 0) {
            for (int i = 0; i < K5; ++i) gr[i] = A[i * K5 + i];
            return;
        }
        const double b0 = th[2] + p[2], b01 = th[3] + p[3], b1 = th[4] + p[4];
        gr[2] = -b1;
        gr[3] = 2.0 * b01;
        gr[4] = -b0;
    }
    bool feasible(const double* p, double tol) const {
        for (int i = 0; i < K5; ++i)
            if (p[i] < lb[i] - tol || p[i] > ub[i] + tol) return false;
        return cons(0, p) <= tol * (1.0 + std::fabs(delta)) && cons(1, p) <= tol;
    }
};

// Augmented Lagrangian value and gradient (box handled by projection).
double al_value(const StepProblem& P, const double* p, const double* lam, double mu, double* grad) {
    double v = quad(P.g, P.A, p);
    for (int i = 0; i < K5; ++i) {
        double ap = 0.0;
        for (int j = 0; j < K5; ++j) ap += P.A[i * K5 + j] * p[j];
        grad[i] = -(P.g[i] - ap);
    }
    for (int j = 0; j < 2; ++j) {
        const double t = P.cons(j, p) + lam[j] / mu;
        if (t > 0.0) {
            v += 0.5 * mu * t * t - lam[j] * lam[j] / (2.0 * mu);
            double cg[K5];
            P.cons_grad(j, p, cg);
            for (int i = 0; i < K5; ++i) grad[i] += mu * t * cg[i];
        } else {
            v -= lam[j] * lam[j] / (2.0 * mu);
        }
    }
    return v;
}

void project(const StepProblem& P, double* p) {
    for (int i = 0; i < K5; ++i) p[i] = std::fmin(P.ub[i], std::fmax(P.lb[i], p[i]));
}

// Hessian of the augmented Lagrangian at p (constraints with c_j + lam_j / mu > 0).
void al_hessian(const StepProblem& P, const double* p, const double* lam, double mu, double* H) {
    std::memcpy(H, P.A, sizeof(double) * K5 * K5);
    for (int j = 0; j < 2; ++j) {
        const double t = P.cons(j, p) + lam[j] / mu;
        if (t <= 0.0) continue;
        double cg[K5];
        P.cons_grad(j, p, cg);
        for (int a = 0; a < K5; ++a)
            for (int b = 0; b < K5; ++b) H[a * K5 + b] += mu * cg[a] * cg[b];
        if (j == 1) {  // second derivatives of (b01^2 - b0 b1)
            H[3 * K5 + 3] += mu * t * 2.0;
            H[2 * K5 + 4] -= mu * t;
            H[4 * K5 + 2] -= mu * t;
        }
    }
}

// Solve H_FF d_F = -g_F on the free set by Cholesky, adding a Levenberg shift until
// H_FF + shift I is positive definite.
bool newton_dir(const double* H, const double* g, const bool* fr, double* d) {
    int idx[K5], m = 0;
    for (int i = 0; i < K5; ++i) {
        d[i] = 0.0;
        if (fr[i]) idx[m++] = i;
    }
    if (m == 0) return false;
    double scale = 0.0;
    for (int a = 0; a < m; ++a) scale = std::fmax(scale, std::fabs(H[idx[a] * K5 + idx[a]]));
    double shift = 0.0;
    for (int attempt = 0; attempt < 40; ++attempt) {
        double L[K5][K5] = {};
        bool ok = true;
        for (int a = 0; a < m && ok; ++a)
            for (int b = 0; b <= a; ++b) {
                double s = H[idx[a] * K5 + idx[b]] + (a == b ? shift : 0.0);
                for (int q = 0; q < b; ++q) s -= L[a][q] * L[b][q];
                if (a == b) {
                    if (s <= 0.0) { ok = false; break; }
                    L[a][a] = std::sqrt(s);
                } else {
                    L[a][b] = s / L[b][b];
                }
            }
        if (ok) {
            double y[K5];
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
    double lam[2] = {0.0, 0.0}, mu = 10.0;
    project(P, p);
    double prev_viol = 1e300;
    for (int outer = 0; outer < 60; ++outer) {
        // projected Newton on the box-constrained augmented Lagrangian
        double gr[K5];
        double f = al_value(P, p, lam, mu, gr);
        for (int it = 0; it < 500; ++it) {
            bool fr[K5];
            double pg = 0.0;
            for (int i = 0; i < K5; ++i) {
                const bool at_lo = p[i] <= P.lb[i] + 1e-14 && gr[i] > 0.0;
                const bool at_hi = p[i] >= P.ub[i] - 1e-14 && gr[i] < 0.0;
                fr[i] = !(at_lo || at_hi);
                if (fr[i]) pg = std::fmax(pg, std::fabs(gr[i]));
            }
            if (pg < 1e-13) break;
            double H[K5 * K5], d[K5];
            al_hessian(P, p, lam, mu, H);
            if (!newton_dir(H, gr, fr, d)) break;
            double a = 1.0, pn[K5], grn[K5], fn = f;
            bool ok = false;
            for (int bt = 0; bt < 60; ++bt) {
                for (int i = 0; i < K5; ++i) pn[i] = p[i] + a * d[i];
                project(P, pn);
                fn = al_value(P, pn, lam, mu, grn);
                double dec = 0.0;
                for (int i = 0; i < K5; ++i) dec += gr[i] * (p[i] - pn[i]);
                if (fn <= f - 1e-4 * std::fabs(dec) || fn < f) {
                    ok = true;
                    break;
                }
                a *= 0.5;
            }
            if (!ok) break;
            double step = 0.0;
            for (int i = 0; i < K5; ++i) step = std::fmax(step, std::fabs(pn[i] - p[i]));
            std::memcpy(p, pn, sizeof(pn));
            std::memcpy(gr, grn, sizeof(grn));
            f = fn;
            if (step < 1e-16) break;
        }
        // stop at a KKT point: feasible and multipliers no longer moving (complementary
        // slackness), otherwise update the multipliers and, if feasibility stalls, mu
        double viol = 0.0, dlam = 0.0, lmax = 0.0;
        for (int j = 0; j < 2; ++j) {
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

bool invert5(const double* a, double* out) {
    double w[K5][2 * K5];
    for (int i = 0; i < K5; ++i)
        for (int j = 0; j < K5; ++j) {
            w[i][j] = a[i * K5 + j];
            w[i][K5 + j] = i == j;
        }
    for (int c = 0; c < K5; ++c) {
        int piv = c;
        for (int r = c + 1; r < K5; ++r)
            if (std::fabs(w[r][c]) > std::fabs(w[piv][c])) piv = r;
        if (w[piv][c] == 0.0) return false;
        for (int j = 0; j < 2 * K5; ++j) std::swap(w[c][j], w[piv][j]);
        const double d = w[c][c];
        for (int j = 0; j < 2 * K5; ++j) w[c][j] /= d;
        for (int r = 0; r < K5; ++r)
            if (r != c) {
                const double f = w[r][c];
                for (int j = 0; j < 2 * K5; ++j) w[r][j] -= f * w[c][j];
            }
    }
    for (int i = 0; i < K5; ++i)
        for (int j = 0; j < K5; ++j) out[i * K5 + j] = w[i][K5 + j];
    return true;
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

uint64_t mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" int globe_aireml_step(const double grad[5], const double ai[25], const double theta[5],
                                 double delta, double step[5]) {
    if (!grad || !ai || !theta || !step) return GLOBE_E_INVALID;
    StepProblem P;
    std::memcpy(P.g, grad, sizeof(P.g));
    std::memcpy(P.A, ai, sizeof(P.A));
    std::memcpy(P.th, theta, sizeof(P.th));
    P.delta = delta;
    for (int i = 0; i < K5; ++i) {
        P.ub[i] = 10.0;
        P.lb[i] = i == 3 ? -10.0 : std::fmax(-10.0, 1e-5 - theta[i]);
    }
    double Ai[K5 * K5], p[K5];
    if (invert5(ai, Ai)) {
        for (int i = 0; i < K5; ++i) {
            p[i] = 0.0;
            for (int j = 0; j < K5; ++j) p[i] += Ai[i * K5 + j] * grad[j];
        }
        bool spd = true;  // the Newton point minimises q only when AI is positive definite
        for (int i = 0; i < K5; ++i) spd &= ai[i * K5 + i] > 0.0;
        if (spd && P.feasible(p, 0.0)) {
            std::memcpy(step, p, sizeof(p));
            return GLOBE_OK;
        }
    }
    // The covariance constraint is not convex, so local minimisers can differ. NLopt
    // starts from the zero step (:1163); the zero step and two points along the
    // (projected) Newton direction are tried here and the best feasible point is kept.
    double starts[4][K5];
    for (int i = 0; i < K5; ++i) {
        starts[0][i] = 0.0;
        starts[1][i] = std::fmin(P.ub[i], std::fmax(P.lb[i], p[i]));
        starts[2][i] = 0.5 * starts[1][i];
        starts[3][i] = 0.1 * starts[1][i];
    }
    bool found = false;
    double best[K5], fbest = 0.0;
    for (int s = 0; s < 4; ++s) {
        double q[K5];
        std::memcpy(q, starts[s], sizeof(q));
        for (int i = 0; i < K5; ++i)
            if (!std::isfinite(q[i])) q[i] = 0.0;
        if (!solve_al(P, q)) continue;
        const double f = quad(P.g, P.A, q);
        if (!found || f < fbest) {
            found = true;
            fbest = f;
            std::memcpy(best, q, sizeof(q));
        }
    }
    if (!found) return GLOBE_E_NUMERIC;
    std::memcpy(step, best, sizeof(best));
    return GLOBE_OK;
}

namespace {

// Probes and trace weights for one gradient evaluation.
//   Gaussian tilde-V (varCompEst.cpp:775-862): y ~ N(0, tilde-V); weights = x0 =
//   tilde-V^-1 y by the per-individual 2x2 Woodbury solve on the host.
//   Rademacher: z in {-1, +1}; weights = z; PCG starts from Q^-1 z.
struct Probes {
    std::vector<double> y, w;
    bool has_x0 = false;
};

Probes make_probes(globe_ctx* c, const double th[5], int nmc, int kind, uint64_t seed) {
    Probes P;
    const int64_t nrec = c->nrec;
    const int32_t n = c->n;
    P.y.resize((size_t)nrec * nmc);
    P.w.resize((size_t)nrec * nmc);
    const double se = th[0], sg = th[1], sb0 = th[2], sb01 = th[3], sb1 = th[4];
    for (int s = 0; s < nmc; ++s) {
        std::mt19937_64 gen(mix(seed ^ (0x1000193ull * (uint64_t)(s + 1))));
        double* ys = P.y.data() + (size_t)s * nrec;
        double* ws = P.w.data() + (size_t)s * nrec;
        if (kind == 1) {
            for (int64_t r = 0; r < nrec; ++r) ys[r] = (gen() >> 63) ? 1.0 : -1.0;
            std::memcpy(ws, ys, sizeof(double) * nrec);
            continue;
        }
        std::normal_distribution<double> nd(0.0, 1.0);
        const double l00 = std::sqrt(sb0), l10 = sb01 / l00;
        double d11 = sb1 - sb01 * sb01 / sb0;
        if (d11 < 0) d11 = 0;  // clipped as compute_chol_2x2 does (:806-810)
        const double l11 = std::sqrt(d11);
        std::vector<double> z1(n), z2(2 * (size_t)n);
        for (auto& v : z1) v = nd(gen);
        for (auto& v : z2) v = nd(gen);
        for (int64_t r = 0; r < nrec; ++r) {
            const int32_t i = c->h_indiv[r];
            const double t = c->h_times[r];
            ys[r] = std::sqrt(sg) * z1[i] + std::sqrt(se) * nd(gen) + z2[2 * i] * (l00 + l10 * t) +
                    z2[2 * i + 1] * l11 * t;
        }
        // tilde-V^-1 y per individual block: Q = se I + U C U^T, C = [[sg + sb0, sb01], [sb01, sb1]]
        for (int32_t b = 0; b < n; ++b) {
            const int64_t r0 = c->h_blk_off[b], r1 = c->h_blk_off[b + 1];
            double a01 = 0, a11 = 0, w0 = 0, w1 = 0;
            for (int64_t q = r0; q < r1; ++q) {
                const double t = c->h_times[q];
                a01 += t;
                a11 += t * t;
                w0 += ys[q];
                w1 += t * ys[q];
            }
            const double a00 = (double)(r1 - r0);
            const double c00 = sg + sb0, c01 = sb01, c11 = sb1;
            const double m00 = se + c00 * a00 + c01 * a01, m01 = c00 * a01 + c01 * a11;
            const double m10 = c01 * a00 + c11 * a01, m11 = se + c01 * a01 + c11 * a11;
            const double v0 = c00 * w0 + c01 * w1, v1 = c01 * w0 + c11 * w1;
            const double det = m00 * m11 - m01 * m10;
            const double k0 = (m11 * v0 - m01 * v1) / det, k1 = (m00 * v1 - m10 * v0) / det;
            for (int64_t q = r0; q < r1; ++q) ws[q] = (ys[q] - k0 - k1 * c->h_times[q]) / se;
        }
    }
    P.has_x0 = kind == 0;
    return P;
}

struct Eval {
    double grad[5], ai[25];
};

void evaluate(globe_ctx* c, const double* y, const double th[5], const globe_aireml_opts& o,
              uint64_t eval_id, bool want_ai, Eval& e) {
    int rc = globe_set_varcomp(c, th);
    if (rc != GLOBE_OK) throw Error(rc, c->err);
    Probes P = make_probes(c, th, o.num_mc, o.probes, mix(o.seed + 0x9E37ull * (eval_id + 1)));
    globe_pcg_opts po = o.pcg;
    rc = globe::aireml_eval_impl(c, y, P.y.data(), P.w.data(), P.has_x0 ? P.w.data() : nullptr,
                                 o.num_mc, &po, want_ai, e.grad, e.ai, nullptr, nullptr);
    if (rc != GLOBE_OK) throw Error(rc, c->err);
}

double norm5(const double* v) {
    double s = 0;
    for (int i = 0; i < K5; ++i) s += v[i] * v[i];
    return std::sqrt(s);
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
    try {
        const globe_aireml_opts& o = *opts;
        Logger log{logfn, user};
        globe::require(o.num_mc >= 1 && o.num_mc <= kMaxCols, GLOBE_E_INVALID,
                       "numMC must be in [1, 64]");
        double th[5];
        std::memcpy(th, theta, sizeof(th));
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
        log("Initial variance components:  sig2e = %g, sig2g = %g, sig2b0 = %g, sigb01 = %g, sig2b1 = %g",
            th[0], th[1], th[2], th[3], th[4]);
        for (int it = 0; it < o.max_iter; ++it) {
            log("---------- Iteration %d ----------", it + 1);
            const auto t0 = std::chrono::steady_clock::now();
            if (it == 0) evaluate(c, y, th, o, eval_id++, true, cur);
            if (it == 1) {
                gnt = 0.1 * norm5(cur.grad);
                log("Gradient norm threshold: %g", gnt);
            }
            if (num_rej > 5) {
                // random restart in [0.1, 0.9] (varCompEst.cpp:1431-1451), seeded here
                log("Too many rejections at current point. Re-initalizing optimization.");
                std::uniform_real_distribution<double> u(0.0, 1.0);
                for (int j = 0; j < 5; ++j) th[j] = 0.1 + 0.8 * u(restart_rng);
                if (th[3] > std::sqrt(th[2] * th[4])) th[3] = std::sqrt(th[2] * th[4]) - 1e-3;
                evaluate(c, y, th, o, eval_id++, true, cur);
                num_rej = 0;
            }
            const double th_old[5] = {th[0], th[1], th[2], th[3], th[4]};
            // ---- oneIter_AIREML (varCompEst.cpp:1148-1344)
            double p[5];
            int rc = globe_aireml_step(cur.grad, cur.ai, th, Delta, p);
            if (rc != GLOBE_OK) {
                log("Trust-region step failed. Reducing trust region radius.");
                Delta *= 0.75;
                dangerous = true;
                num_rej += 1;
            } else {
                log("Predicted step: %g %g %g %g %g", p[0], p[1], p[2], p[3], p[4]);
                double scale = 1.0;
                bool accepted = false, converged_pred = false;
                int red = 0;
                while (!accepted && red <= 10) {
                    double sp[5], tn[5];
                    for (int j = 0; j < 5; ++j) {
                        sp[j] = scale * p[j];
                        tn[j] = th[j] + sp[j];
                    }
                    bool violated = false;
                    for (int j = 0; j < 5; ++j) {
                        if (j != 3 && tn[j] < 1e-6) {
                            tn[j] = 1e-6;
                            violated = true;
                        }
                        if (tn[j] > 10.0) {
                            tn[j] = 10.0;
                            violated = true;
                        }
                    }
                    if (tn[3] > std::sqrt(tn[2] * tn[4])) {
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
                    evaluate(c, y, tn, o, eval_id++, false, nw);
                    double dll = 0.0, pred = 0.0, spa = 0.0;
                    for (int a = 0; a < 5; ++a) {
                        dll += sp[a] * 0.5 * (cur.grad[a] + nw.grad[a]);
                        pred += cur.grad[a] * sp[a];
                        for (int b = 0; b < 5; ++b) spa += sp[a] * cur.ai[a * 5 + b] * sp[b];
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
                    if (grad_check && norm5(cur.grad) > gnt && norm5(nw.grad) > 2 * norm5(cur.grad)) {
                        log("Dangerous model deviation detected at scale %g. Reducing step size.", scale);
                        scale *= 0.5;
                        red += 1;
                        continue;
                    }
                    const double rho = dll / pred;
                    log("Rho: %g", rho);
                    double npd = 0.0;
                    for (int j = 0; j < 5; ++j) npd += std::fabs(cur.ai[j * 5 + j] * sp[j]);
                    if (rho < 0.25) Delta = 0.25 * npd;
                    else if (rho > 0.99) Delta = std::fmax(Delta, 3.5 * npd);
                    if (rho > 0.25) {
                        std::memcpy(th, tn, sizeof(th));
                        std::memcpy(cur.grad, nw.grad, sizeof(cur.grad));
                        // AI at the accepted point from the trial's fit and f_theta
                        // (computeDerivative_AImat, :1321-1330)
                        globe_pcg_opts po = o.pcg;
                        const int arc = globe::aireml_eval_impl(c, nullptr, nullptr, nullptr, nullptr, 0, &po,
                                                                true, nullptr, cur.ai, nullptr, nullptr);
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
            log("Estimated variance components:  sig2e = %g, sig2g = %g, sig2b0 = %g, sigb01 = %g, sig2b1 = %g",
                th[0], th[1], th[2], th[3], th[4]);
            const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            total_secs += secs;
            it_done = it + 1;
            log("Iteration %d took %g seconds.", it + 1, secs);
            double dn = 0.0, on = 0.0;
            for (int j = 0; j < 5; ++j) {
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
        // SEs from the AI matrix at the final point (:1484-1491)
        double Ai[25];
        const bool inv_ok = invert5(cur.ai, Ai);
        const char* nm[5] = {"sig2e", "sig2g", "sig2b0", "sigb01", "sig2b1"};
        log("Final variance components +/- SEs:");
        for (int j = 0; j < 5; ++j) {
            const double s = inv_ok ? std::sqrt(Ai[j * 5 + j]) : NAN;
            if (se_out) se_out[j] = s;
            log("  %6s = %g +/- %g", nm[j], th[j], s);
        }
        std::memcpy(theta, th, sizeof(th));
        if (ai_out) std::memcpy(ai_out, cur.ai, sizeof(cur.ai));
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
