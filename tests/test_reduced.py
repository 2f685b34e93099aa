"""Reduced model (varCompEst_reducedModel_noTime.cpp): V = sig2g A + sig2e I, one record
per individual, two components, diagonal preconditioner, no covariance constraint.

CPU tests pin the oracle against dense closed forms built from the reduced model's own
definition (computeV_ij :288-296, P, f_theta :524-548, the trace estimator :592-660),
the exact-trace replay of the AI-REML loop against the dense REML optimum, the step
solver against the Newton step and SLSQP, and the CLI's input checks. GPU tests compare
the device path with the oracle and the replay.
"""
import os
import subprocess

import numpy as np
import pytest
from scipy.optimize import minimize

import globe_amd
import oracle as O
from globe_amd import synth
from helpers import rel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "globe_amd", "bin", "runAIREML_reducedModel")


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def case():
    # seed chosen so that the REML optimum is interior (with 56 individuals it often
    # lies on the sig2e or sig2g boundary)
    return synth.make_reduced_case(n=56, m=400, seed=33, theta=(0.5, 0.5))


def oracle_for(d, theta2):
    n = d.n
    return O.Oracle(np.arange(n, dtype=np.int32), np.zeros(n), d.covars,
                    (theta2[0], theta2[1], 0.0, 0.0, 0.0), grm_lt=d.grm_lt, n=n)


def dense_terms(d, theta2):
    """V, P and dV of the reduced model from its definition."""
    A = synth.grm_dense(d.grm_lt, d.n)
    se, sg = theta2
    V = sg * A + se * np.eye(d.n)
    Vi = np.linalg.inv(V)
    X = d.covars
    P = Vi - Vi @ X @ np.linalg.solve(X.T @ Vi @ X, X.T @ Vi)
    return A, V, P, [np.eye(d.n), A]


def reml_loglik(theta2, d):
    _, V, P, _ = dense_terms(d, theta2)
    X = d.covars
    Vi = np.linalg.inv(V)
    return -0.5 * (np.linalg.slogdet(V)[1] + np.linalg.slogdet(X.T @ Vi @ X)[1] + d.y @ P @ d.y)


def reml_optimum(d):
    r = minimize(lambda t: -reml_loglik(t, d), np.array([0.5, 0.5]), method="L-BFGS-B",
                 bounds=[(1e-6, 10.0)] * 2, options={"ftol": 1e-15, "gtol": 1e-12, "maxiter": 2000})
    return r.x


# ------------------------------------------------------------------------------ CPU
def test_oracle_terms_match_dense_reduced_model(case):
    d = case
    th = (0.45, 0.55)
    orc = oracle_for(d, th)
    A, V, P, dV = dense_terms(d, th)
    Py = P @ d.y
    f = np.column_stack([D @ Py for D in dV])
    rng = np.random.default_rng(3)
    probes = rng.choice([-1.0, 1.0], size=(d.n, 9))
    out = orc.aireml_eval_reduced(d.y, probes, probes, tol=1e-13)
    assert rel(out["first"], Py @ f) < 1e-9
    tr = np.array([np.mean([(D @ P @ probes[:, s]) @ probes[:, s] for s in range(9)]) for D in dV])
    assert rel(out["trace"], tr) < 1e-9
    assert rel(out["grad"], 0.5 * (Py @ f - tr)) < 1e-9
    assert rel(out["ai"], 0.5 * f.T @ P @ f) < 1e-9
    # unit probes sqrt(N) e_s: the estimator is the exact trace tr(P dV)
    E = np.sqrt(d.n) * np.eye(d.n)
    ex = orc.aireml_eval_reduced(d.y, E, E, tol=1e-13)
    assert rel(ex["trace"], [np.trace(P @ D) for D in dV]) < 1e-9
    # the reference's probes: y ~ N(0, (sig2g + sig2e) I) (sample_Yvs :563-573), weights
    # tilde-V^-1 y = y / (sig2g + sig2e) (:592-602)
    yv = rng.standard_normal((d.n, 3))
    wv = yv / (th[0] + th[1])
    g = orc.aireml_eval_reduced(d.y, yv, wv, wv, tol=1e-13)
    trg = np.array([np.mean([(D @ P @ yv[:, s]) @ wv[:, s] for s in range(3)]) for D in dV])
    assert rel(g["trace"], trg) < 1e-9


def test_exact_trace_replay_reaches_reml_optimum(case):
    d = case
    opt = reml_optimum(d)
    th, traj, iters = O.aireml_fit_reduced(oracle_for(d, (0.5, 0.5)), d.y, (0.5, 0.5), tol=1e-6)
    assert 1 <= iters < 50
    # the loop stops once the predicted gain g^T AI^-1 g / 2 is below 1e-2 (:920-925),
    # so theta is within that model step of the optimum, measured in the AI metric, and
    # the REML log-likelihood within about that gain of its maximum
    assert np.all(opt > 1e-3)
    assert 0.0 <= reml_loglik(opt, d) - reml_loglik(th, d) < 0.02
    orc = oracle_for(d, th)
    E = np.sqrt(d.n) * np.eye(d.n)
    ev = orc.aireml_eval_reduced(d.y, E, E, tol=1e-13)
    dist = float(np.sqrt((th - opt) @ ev["ai"] @ (th - opt)))
    assert dist < 0.2, (th, opt, dist)
    # and at the optimum the exact REML score vanishes
    orc.set_theta((opt[0], opt[1], 0, 0, 0))
    eo = orc.aireml_eval_reduced(d.y, E, E, tol=1e-13)
    assert np.abs(eo["grad"]).max() < 1e-4 * np.abs(ev["ai"]).max()


def _step(g, A, th, delta):
    L = globe_amd.load_library()
    p = np.zeros(2)
    g, A, th = (np.ascontiguousarray(v, dtype=np.float64) for v in (g, A, th))
    rc = L.globe_aireml_step_reduced(g.ctypes.data, A.ctypes.data, th.ctypes.data, delta, p.ctypes.data)
    return rc, p


def test_reduced_step_newton_and_constrained():
    rng = np.random.default_rng(5)
    for trial in range(16):
        M = rng.standard_normal((2, 2))
        A = 40.0 * (M @ M.T + 0.3 * np.eye(2))
        th = rng.uniform(0.1, 0.9, size=2)
        g = A @ (rng.standard_normal(2) * (0.05 if trial < 4 else 0.8))
        delta = 1e20 if trial % 2 == 0 else float(abs(np.diag(A) @ np.linalg.solve(A, g)) * 0.3 + 0.1)
        rc, p = _step(g, A, th, delta)
        assert rc == 0
        lb = np.maximum(-10.0, 1e-5 - th)
        pn = np.linalg.solve(A, g)
        if np.all(pn >= lb) and np.all(pn <= 10) and np.diag(A) @ pn <= delta:
            assert np.allclose(p, pn, rtol=1e-12, atol=1e-14)
            continue
        # convex QP: the minimiser is unique; SLSQP finds it too
        cons = [{"type": "ineq", "fun": lambda q: delta - np.diag(A) @ q}]
        r = minimize(lambda q: -(g @ q - 0.5 * q @ A @ q), np.zeros(2), method="SLSQP",
                     bounds=list(zip(lb, [10.0, 10.0])), constraints=cons,
                     options={"ftol": 1e-15, "maxiter": 1000})
        assert np.all(p >= lb - 1e-9) and np.diag(A) @ p <= delta + 1e-9 * (1 + abs(delta))
        f = lambda q: -(g @ q - 0.5 * q @ A @ q)  # noqa: E731
        assert f(p) <= f(r.x) + 1e-7 * (1 + abs(f(r.x)))
        assert np.allclose(p, r.x, atol=1e-5 * (1 + np.abs(r.x).max()))


def _run(args):
    return subprocess.run(args, capture_output=True, text=True, timeout=300)


def test_reduced_cli_input_errors(case, tmp_path):
    d = case
    assert _run([CLI, "a", "b"]).returncode == 1
    inp = str(tmp_path / "in.txt")
    synth.write_reduced_input(inp, d)
    grm = str(tmp_path / "g")
    with open(grm + ".grm.bin", "wb") as fh:
        fh.write(bytes(16))
    r = _run([CLI, inp, grm, str(tmp_path / "o")])
    assert r.returncode == 1 and "Error reading" in r.stderr
    dup = str(tmp_path / "dup.txt")
    lines = open(inp).read().splitlines()
    open(dup, "w").write("\n".join(lines + [lines[1]]) + "\n")
    r = _run([CLI, dup, grm, str(tmp_path / "o")])
    assert r.returncode == 1 and "appears twice" in r.stderr
    nohdr = str(tmp_path / "nohdr.txt")
    open(nohdr, "w").write("ID PHENO x\n1 0.5 1.0\n")
    r = _run([CLI, nohdr, grm, str(tmp_path / "o")])
    assert r.returncode == 1 and "Missing required columns" in r.stderr
    nocov = str(tmp_path / "nocov.txt")
    open(nocov, "w").write("INDIV PHENO\n1 0.5\n")
    r = _run([CLI, nocov, grm, str(tmp_path / "o")])
    assert r.returncode == 1 and "covariate" in r.stderr


# ------------------------------------------------------------------------------ GPU
def device_for(d, factory, theta2):
    g = factory()
    n = d.n
    g.set_cohort(np.arange(n, dtype=np.int32), np.zeros(n), d.covars)
    g.readFullGRM(d.grm_lt)
    g.set_varcomp((theta2[0], theta2[1], 0.0, 0.0, 0.0))
    return g


@pytest.mark.gpu
def test_reduced_eval_matches_oracle(case, gpu_ctx_factory):
    d = case
    th = (0.45, 0.55)
    _say("reduced eval", d.n)
    g = device_for(d, gpu_ctx_factory, th)
    orc = oracle_for(d, th)
    rng = np.random.default_rng(8)
    probes = rng.choice([-1.0, 1.0], size=(d.n, 15))
    og = g.aireml_eval(d.y, probes, probes, tol=1e-11, reduced=True)
    oo = orc.aireml_eval_reduced(d.y, probes, probes, tol=1e-11)
    for key in ("first", "trace", "grad", "ai"):
        assert og[key].shape == oo[key].shape
        assert rel(og[key], oo[key]) < 1e-7, key
    yv = rng.standard_normal((d.n, 4))
    wv = yv / (th[0] + th[1])
    og = g.aireml_eval(d.y, yv, wv, wv, tol=1e-11, reduced=True)
    oo = orc.aireml_eval_reduced(d.y, yv, wv, wv, tol=1e-11)
    assert rel(og["grad"], oo["grad"]) < 1e-7 and rel(og["ai"], oo["ai"]) < 1e-7
    g.close()


@pytest.mark.gpu
def test_reduced_eval_rejects_full_model_cohorts(case, gpu_ctx_factory):
    d = case
    _say("reduced eval rejects")
    g = device_for(d, gpu_ctx_factory, (0.5, 0.5))
    g.set_varcomp((0.5, 0.5, 0.1, 0.0, 0.0))
    probes = np.ones((d.n, 1))
    with pytest.raises(globe_amd.GlobeError):
        g.aireml_eval(d.y, probes, probes, reduced=True)
    # two records per individual
    n = d.n // 2
    g.set_cohort(np.repeat(np.arange(n, dtype=np.int32), 2), np.zeros(2 * n), d.covars[: 2 * n])
    g.readFullGRM(d.grm_lt[: n * (n + 1) // 2])
    g.set_varcomp((0.5, 0.5, 0.0, 0.0, 0.0))
    with pytest.raises(globe_amd.GlobeError):
        g.aireml_eval(d.y[: 2 * n], probes[: 2 * n], probes[: 2 * n], reduced=True)
    g.close()


@pytest.mark.gpu
def test_reduced_fit_matches_exact_replay(case, gpu_ctx_factory):
    d = case
    _say("reduced fit exact", d.n)
    g = device_for(d, gpu_ctx_factory, (0.5, 0.5))
    fit = g.aireml_fit(d.y, (0.5, 0.5), max_iter=50, tol=1e-6, num_mc=d.n, probes="exact",
                       pcg_tol=1e-10, reduced=True)
    th, traj, iters = O.aireml_fit_reduced(oracle_for(d, (0.5, 0.5)), d.y, (0.5, 0.5), tol=1e-6,
                                           pcg_tol=1e-10)
    assert fit["iters"] == iters
    assert rel(fit["theta"], th) < 1e-6
    assert fit["ai"].shape == (2, 2) and np.all(np.isfinite(fit["se"]))
    assert any("Iteration 1 took" in s for s in fit["log"])
    g.close()


@pytest.mark.gpu
def test_reduced_cli_runs_exact_and_logs(case, tmp_path):
    d = case
    _say("runAIREML_reducedModel", d.n)
    inp = str(tmp_path / "in.txt")
    synth.write_reduced_input(inp, d)
    grm = str(tmp_path / "g")
    synth.write_grm_bin(grm, d.grm_lt)
    out = str(tmp_path / "o")
    r = _run([CLI, inp, grm, out, "--numMC", str(d.n), "--probes", "exact", "--tol", "1e-6",
              "--pcgTol", "1e-10", "--sig2e", "0.5", "--sig2g", "0.5", "--bogus", "--loglik", "exact"])
    assert r.returncode == 0, r.stderr + r.stdout
    log = open(out + ".log").read()
    assert log.startswith("Command:") and "Final variance components" in log
    th, _, _ = O.aireml_fit_reduced(oracle_for(d, (0.5, 0.5)), d.y, (0.5, 0.5), tol=1e-6,
                                    pcg_tol=1e-10)
    est = [ln for ln in log.splitlines() if ln.startswith("Estimated variance components:")][-1]
    vals = [float(x.split("=")[1]) for x in est.split(":", 1)[1].split(",")]
    assert rel(vals, th) < 1e-5
    # --loglik exact: the REML log-likelihood at the printed estimates (dense closed form
    # plus the -(n - q)/2 log 2 pi constant the build reports)
    ln = next(x for x in log.splitlines() if x.startswith("REML log-likelihood"))
    ll = float(ln.split(":")[1].split("(")[0])
    want = reml_loglik(np.array(vals), d) - 0.5 * (d.n - d.covars.shape[1]) * np.log(2 * np.pi)
    assert abs(ll - want) < 1e-6 * abs(want), (ll, want)
