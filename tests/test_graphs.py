"""PCG iterations replayed from captured HIP graphs against the same solves enqueued
kernel by kernel (calls through the C ABI).

A solve captures one iteration per digit configuration after running it once eagerly,
then replays it; each replay's output feeds the next iteration, so a single wrong replay
changes the solution. Bar: bit-identical solutions, iteration counts and per-iteration
Lanczos coefficients (the kernels are deterministic), over several solves in a row on
one context with host work and synchronisations between them.
"""
import numpy as np
import pytest

import globe_amd as G
from globe_amd import synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _ctx(gpu_ctx_factory, kind):
    if kind == "geno":
        d = synth.make_geno_case(n=1500, m=3000, lam=4, seed=31, miss=0.01)
    else:
        d = synth.make_grm_case(n=1500, m=2000, lam=4, seed=32)
    c = d.cohort
    g = gpu_ctx_factory()
    g.set_cohort(c.indiv, c.times, c.X)
    if kind == "geno":
        g.prepGenotypeInput(d.bed_rows, 0.01, 0.2)
    else:
        g.readFullGRM(d.grm_lt)
    g.set_varcomp(d.theta)
    return g, d


@pytest.mark.parametrize("kind", ["geno", "grm"])
@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_MIXED])
@pytest.mark.parametrize("k", [1, 20])
def test_graph_replay_equals_eager(kind, prec, k, gpu_ctx_factory):
    _say("graphs", kind, prec, k)
    g, d = _ctx(gpu_ctx_factory, kind)
    nrec = d.cohort.nrec
    rng = np.random.default_rng(7)
    B = np.column_stack([d.y, rng.choice([-1.0, 1.0], size=(nrec, k - 1))])[:, :k]
    out = {}
    for on in (False, True, False, True):
        g.set_graphs(on)
        for rule in (G.RULE_PER_COLUMN, G.RULE_MAX_RATIO):
            x, st = g.pcg(B, tol=1e-9, rule=rule, stagnation=False, precision=prec)
            key = (on, rule)
            if key in out:
                # a repeated solve on the same context gives the same bits
                assert np.array_equal(out[key][0], x)
            out[key] = (x, st)
    for rule in (G.RULE_PER_COLUMN, G.RULE_MAX_RATIO):
        xe, se = out[(False, rule)]
        xg, sg = out[(True, rule)]
        assert se["iters"] == sg["iters"] and se["iters"] > 5
        assert np.array_equal(se["col_iters"], sg["col_iters"])
        assert np.array_equal(xe, xg), np.max(np.abs(xe - xg))
    g.close()


@pytest.mark.parametrize("kind", ["geno", "grm"])
def test_graph_replay_lanczos_and_timing(kind, gpu_ctx_factory):
    """Every iteration's alpha and beta (the log-likelihood's Lanczos coefficients) agree
    bit for bit, and timed solves report the same contraction launches with graphs."""
    _say("graphs lanczos", kind)
    g, d = _ctx(gpu_ctx_factory, kind)
    res = {}
    for on in (False, True):
        g.set_graphs(on)
        res[on] = g.reml_loglik(d.y, probes="rademacher", nprobe=24, seed=5, tol=1e-8,
                                precision=G.PREC_FP64)
    for key in ("loglik", "logdet_v", "ypy", "max_steps"):
        assert res[False][key] == res[True][key], key
    which = (0, 1) if kind == "geno" else (2,)
    nrec = d.cohort.nrec
    B = np.column_stack([d.y, np.random.default_rng(3).choice([-1.0, 1.0], size=(nrec, 7))])
    launches = {}
    for on in (False, True):
        g.set_graphs(on)
        g.pcg(B, tol=1e-8, rule=G.RULE_PER_COLUMN, stagnation=False, precision=G.PREC_MIXED)
        g.profile(True)
        _, st = g.pcg(B, tol=1e-8, rule=G.RULE_PER_COLUMN, stagnation=False, precision=G.PREC_MIXED)
        ks = [g.kernel_stats(w) for w in which] + [g.kernel_stats(3)]
        g.profile(False)
        launches[on] = [s["launches"] for s in ks]
        for s in ks:
            if s["launches"]:
                assert s["ms_total"] > 0.0
        assert sum(launches[on]) >= st["iters"]
    assert launches[False] == launches[True]
    g.close()
