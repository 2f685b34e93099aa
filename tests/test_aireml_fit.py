"""The full-model AI-REML fit (runAIREML, varCompEst.cpp:1347-1491).

CPU: the oracle's replay of the loop (oracle.aireml_fit_replay) with exact trace terms
(unit probes) reaches the optimum of the dense REML log-likelihood of the same model,
built from its definition (V = sig2g J G J^T + per-individual intercept/slope blocks +
sig2e I). GPU: the device fit with Rademacher probes equals the replay fed the same
probes (the oracle's restatement of the device generator, keyed per evaluation like the
device) and the same step solver: variance components within 1e-6 and the same
iteration count (identical probe vectors, as the reference's own parity statement asks).
"""
import ctypes as C

import numpy as np
import pytest
from scipy.optimize import minimize

import globe_amd
import oracle as O
from globe_amd import synth
from helpers import dense_V


def _say(*a):
    print("[case]", *a, flush=True)


def step_fn(g, A, th, delta):
    L = globe_amd.load_library()
    L.globe_aireml_step.argtypes = [C.c_void_p] * 3 + [C.c_double, C.c_void_p]
    p = np.zeros(5)
    g = np.ascontiguousarray(g, dtype=np.float64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    th = np.ascontiguousarray(th, dtype=np.float64)
    rc = L.globe_aireml_step(g.ctypes.data, A.ctypes.data, th.ctypes.data, float(delta), p.ctypes.data)
    return p if rc == 0 else None


def probe_seed(seed, e):
    L = globe_amd.load_library()
    L.globe_probe_seed.argtypes = [C.c_uint64, C.c_uint64]
    L.globe_probe_seed.restype = C.c_uint64
    return int(L.globe_probe_seed(seed, e))


def make_case(n, lam, m, seed, theta=(0.6, 0.4, 0.4, 0.1, 0.2)):
    rng = np.random.default_rng(seed)
    c = synth.simulate_cohort(n, lam, rng)
    g = synth.make_reduced_case(n=n, m=m, seed=seed + 1)
    G = synth.grm_dense(g.grm_lt, n)
    V, _ = dense_V(c, G, theta)
    beta = np.array([1.0, 0.3, 0.01, 0.05])[: c.X.shape[1]]
    y = c.X @ beta + np.linalg.cholesky(V) @ rng.standard_normal(c.nrec)
    return c, g.grm_lt, G, y


def reml_loglik(theta, c, G, y):
    V, _ = dense_V(c, G, theta)
    X = c.X
    Vi = np.linalg.inv(V)
    XtViX = X.T @ Vi @ X
    P = Vi - Vi @ X @ np.linalg.solve(XtViX, X.T @ Vi)
    s1, ld1 = np.linalg.slogdet(V)
    s2, ld2 = np.linalg.slogdet(XtViX)
    if s1 <= 0 or s2 <= 0:
        return -np.inf
    return -0.5 * (ld1 + ld2 + y @ P @ y)


# ------------------------------------------------------------------------------ CPU
def test_replay_reaches_dense_reml_optimum():
    # sig2g and sig2b0 are both per-individual effects, told apart only through the GRM's
    # off-diagonal; this case (3,000 SNPs, sig2g 1.0) has an interior optimum
    c, lt, G, y = make_case(n=120, lam=2.0, m=3000, seed=43, theta=(0.5, 1.0, 0.4, 0.1, 0.2))
    orc = O.Oracle(c.indiv, c.times, c.X, (0.5, 0.5, 0.3, 0.0, 0.3), grm_lt=lt, n=c.n)
    unit = np.sqrt(c.nrec) * np.eye(c.nrec)
    th, traj, its = O.aireml_fit_replay(orc, y, (0.5, 0.5, 0.3, 0.0, 0.3), lambda e, t: (unit, unit, None), step_fn,
                                        max_iter=60, tol=1e-7, pcg_tol=1e-12)
    cons = [{"type": "ineq", "fun": lambda t: t[2] * t[4] - t[3] ** 2}]
    r = minimize(lambda t: -reml_loglik(t, c, G, y), th, method="SLSQP",
                 bounds=[(1e-6, 10.0), (1e-6, 10.0), (1e-6, 10.0), (-10.0, 10.0), (1e-6, 10.0)],
                 constraints=cons, options={"ftol": 1e-14, "maxiter": 500})
    best = max(-r.fun, reml_loglik(r.x, c, G, y))
    # the loop stops within its stop rule's reach of the optimum
    assert reml_loglik(th, c, G, y) >= best - 0.02
    # and ends where the Newton model of the exact score predicts less than the stop
    # rule's 0.01 of further gain (the reference's convergence test, :1221-1226)
    orc.set_theta(tuple(th))
    e = orc.aireml_eval(y, unit, unit, tol=1e-12)
    assert 0.5 * e["grad"] @ np.linalg.solve(e["ai"], e["grad"]) < 0.02
    assert 3 <= its < 60


# ------------------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["rademacher", "gaussian"])
def test_device_fit_equals_replay_with_same_probes(kind, gpu_ctx_factory):
    """Rademacher probes, and the reference's Gaussian tilde-V probes (drawn at each
    evaluation's components, weights and PCG start tilde-V^-1 probe)."""
    _say("aireml-fit-replay", kind)
    c, lt, G, y = make_case(n=300, lam=2.0, m=6000, seed=47, theta=(0.5, 1.0, 0.4, 0.1, 0.2))
    th0 = (0.5, 0.5, 0.3, 0.0, 0.3)
    seed, nmc = 20261203, 30
    g = gpu_ctx_factory()
    g.set_cohort(c.indiv, c.times, c.X)
    g.readFullGRM(lt)
    g.set_varcomp(th0)
    fit = g.aireml_fit(y, th0, max_iter=10, tol=1e-3, num_mc=nmc, probes=kind, seed=seed, pcg_tol=1e-10)
    orc = O.Oracle(c.indiv, c.times, c.X, th0, grm_lt=lt, n=c.n)

    def probes_of(e, t):
        if kind == "rademacher":
            Y = O.probes(1, c.indiv, c.times, t, nmc, probe_seed(seed, e))
            return Y, Y, None
        Y = O.probes(0, c.indiv, c.times, t, nmc, probe_seed(seed, e))
        W = orc.tildeV_solve(Y)  # orc is at t while evaluate() draws
        return Y, W, W

    th, traj, its = O.aireml_fit_replay(orc, y, th0, probes_of, step_fn, max_iter=10, tol=1e-3, pcg_tol=1e-10)
    assert fit["iters"] == its
    assert np.max(np.abs(fit["theta"] - th) / np.abs(th)) < 1e-6, (fit["theta"], th)


@pytest.mark.gpu
def test_mixed_fit_within_1e4_of_fp64_on_genotypes(gpu_ctx_factory):
    """Packed-genotype cohort: the mixed-precision fit (int8 digits relaxed per column)
    and the FP64 fit from the same start and probes agree to 1e-4 on the variance
    components."""
    _say("aireml-fit-mixed-geno")
    d = synth.make_geno_case(n=400, m=2000, lam=2, seed=49, miss=0.01)
    c = d.cohort
    th0 = (0.5, 0.5, 0.3, 0.0, 0.3)
    out = {}
    for prec in (globe_amd.PREC_FP64, globe_amd.PREC_MIXED):
        g = gpu_ctx_factory()
        g.set_cohort(c.indiv, c.times, c.X)
        g.prepGenotypeInput(d.bed_rows, 0.01, 0.2)
        g.set_varcomp(th0)
        out[prec] = g.aireml_fit(d.y, th0, max_iter=6, tol=1e-3, num_mc=15, probes="rademacher", seed=7,
                                 pcg_tol=1e-8, precision=prec)
    t64, tm = out[globe_amd.PREC_FP64]["theta"], out[globe_amd.PREC_MIXED]["theta"]
    assert out[globe_amd.PREC_FP64]["iters"] == out[globe_amd.PREC_MIXED]["iters"]
    # relative, with components near 0 (the covariance sigb01) measured against 0.01
    assert np.max(np.abs(tm - t64) / np.maximum(np.abs(t64), 1e-2)) < 1e-4, (t64, tm)
