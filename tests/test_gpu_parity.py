"""HIP path vs the FP64 oracle on the same seeded inputs (calls through the C ABI).

Bars: bit-exact for decode / counts / QC / staged rows; <= 1e-12 relative for the
FP64 operator pieces; <= 1e-7 relative for FP64 solutions at a PCG tolerance of
1e-9 (the oracle runs the same control flow); mixed precision within 1e-6 of the
FP64 solution.
"""
import numpy as np
import pytest

import globe_amd as G
import oracle as O
from globe_amd import synth
from helpers import golden_oracle, load_golden, rel

pytestmark = pytest.mark.gpu

GENO = ["geno_n203", "geno_n150_odd", "geno_qc_n257"]


def _say(*a):
    print("[case]", *a, flush=True)


def device_for(d, ctx_factory, theta=None):
    g = ctx_factory()
    c = d["cohort"]
    g.set_cohort(c.indiv, c.times, c.X)
    if "grm_lt" in d:
        g.readFullGRM(d["grm_lt"])
    else:
        g.prepGenotypeInput(d["bed_rows"], float(d["min_maf"]), float(d["max_miss"]))
    g.set_varcomp(tuple(d["theta"]) if theta is None else theta)
    return g


@pytest.mark.parametrize("name", GENO)
def test_qc_and_staging_bit_exact(name, gpu_ctx_factory):
    _say("qc", name)
    d = load_golden(name)
    g = device_for(d, gpu_ctx_factory)
    n = d["cohort"].n
    keep, maf, _, alt, miss = synth.qc_reference(d["bed_rows"], n, float(d["min_maf"]), float(d["max_miss"]))
    q = g.qc()
    assert np.array_equal(q["alt"], alt) and np.array_equal(q["miss"], miss)
    assert np.array_equal(q["keep"], keep) and np.array_equal(q["maf"], maf)
    assert g.m == int(d["ref_kept"]) == int(keep.sum())
    assert np.array_equal(g.allele_freqs(), maf[keep])
    rows = g.geno_rows()
    nb = (n + 3) // 4
    kept_rows = d["bed_rows"][keep]
    # the device holds dosages, so a missing call (01) reads back as 11 (dosage 0)
    expect = synth.unpack_codes(kept_rows, n)
    expect[expect == 1] = 3
    assert np.array_equal(synth.unpack_codes(rows[:, :nb], n), expect)
    # phantom pairs and padding hold dosage 0 (code 11)
    codes = synth.unpack_codes(rows, rows.shape[1] * 4)
    assert (codes[:, n:] == 3).all()


@pytest.mark.parametrize("n", [1, 3, 5, 130])
def test_qc_tiny_and_ragged(n, gpu_ctx_factory):
    _say("qc-ragged", n)
    rng = np.random.default_rng(n)
    c = synth.simulate_cohort(n, 1.0, rng)
    rows = rng.integers(0, 256, size=(29, (n + 3) // 4), dtype=np.uint8)
    g = gpu_ctx_factory()
    g.set_cohort(c.indiv, c.times, c.X)
    g.prepGenotypeInput(rows, 0.0, 1.0)
    keep, maf, _, alt, miss = synth.qc_reference(rows, n, 0.0, 1.0)
    q = g.qc()
    assert np.array_equal(q["alt"], alt) and np.array_equal(q["miss"], miss)
    assert np.array_equal(q["keep"], keep)


@pytest.mark.parametrize("name", GENO + ["grm_n180"])
def test_genetic_diag(name, gpu_ctx_factory):
    _say("gdiag", name)
    d = load_golden(name)
    g = device_for(d, gpu_ctx_factory)
    assert rel(g.computeGeneticDiagonal(), golden_oracle(d).genetic_diag()) < 1e-13


@pytest.mark.parametrize("name", GENO)
def test_geno_Wa_matches_oracle(name, gpu_ctx_factory):
    """globe_geno_Wa (the synthetic phenotypes' genetic values) against the oracle's W a."""
    _say("geno_Wa", name)
    d = load_golden(name)
    g = device_for(d, gpu_ctx_factory)
    a = np.random.default_rng(12).standard_normal(g.m)
    assert rel(g.geno_Wa(a), golden_oracle(d).geno_Wa(a)) < 1e-12


@pytest.mark.parametrize("name", GENO + ["grm_n180"])
@pytest.mark.parametrize("prec", [G.PREC_FP64, G.PREC_FP64_VALU])
def test_vx_fp64(name, prec, gpu_ctx_factory):
    _say("vx", name, prec)
    d = load_golden(name)
    g = device_for(d, gpu_ctx_factory)
    orc = golden_oracle(d)
    c = d["cohort"]
    x = np.random.default_rng(3).standard_normal((c.nrec, 5))
    assert rel(g.computeVx_multi(x, precision=prec), orc.vx(x)) < 1e-12
    assert rel(g.computeVx_multi(x, multGRM_uhat=True, precision=prec), orc.vx(x, grm_only=True)) < 1e-12
    x1 = x[:, :1] * 1e3 + 7.0  # large mean: centering cancels
    assert rel(g.computeVx_multi(x1, precision=prec), orc.vx(x1)) < 1e-12


@pytest.mark.parametrize("k", [1, 3, 8, 32, 40])
def test_vx_mixed_and_wide(k, gpu_ctx_factory):
    _say("vx-mixed", k)
    d = load_golden("geno_n203")
    g = device_for(d, gpu_ctx_factory)
    orc = golden_oracle(d)
    x = np.random.default_rng(k).standard_normal((d["cohort"].nrec, k))
    ref = orc.vx(x)
    assert rel(g.computeVx_multi(x, precision=G.PREC_FP64), ref) < 1e-12
    e = rel(g.computeVx_multi(x, precision=G.PREC_MIXED), ref)
    assert e < 1e-7, e


def test_precond_matches_dense_inverse(gpu_ctx_factory):
    _say("precond")
    for name in ["geno_n150_odd", "grm_n180"]:
        d = load_golden(name)
        g = device_for(d, gpu_ctx_factory)
        orc = golden_oracle(d)
        r = np.random.default_rng(1).standard_normal((d["cohort"].nrec, 4))
        assert rel(g.applyPreconditioner(r), orc.precond_apply(orc.precond(), r)) < 1e-11


@pytest.mark.parametrize("rule", [G.RULE_MAX_RATIO, G.RULE_ALL_COLUMNS, G.RULE_PER_COLUMN])
@pytest.mark.parametrize("x0_mode", [G.X0_ZERO, G.X0_PRECOND])
def test_pcg_matches_oracle(rule, x0_mode, gpu_ctx_factory):
    _say("pcg", rule, x0_mode)
    d = load_golden("geno_n203")
    g = device_for(d, gpu_ctx_factory)
    orc = golden_oracle(d)
    c = d["cohort"]
    b = np.column_stack([c.X, d["y"]])
    xg, sg = g.pcg(b, tol=1e-9, rule=rule, x0_mode=x0_mode, stagnation=False)
    xo, so = orc.pcg(b, tol=1e-9, rule=rule, x0_mode=x0_mode, stagnation=False)
    assert rel(xg, xo) < 1e-7
    assert abs(sg["iters"] - so["iters"]) <= 1
    # the reference's own tolerance and stall rule: same iteration count
    xg, sg = g.runPCG_multi(b, tol=1e-5)
    xo, so = orc.pcg(b, tol=1e-5, rule=0, stagnation=True)
    assert sg["iters"] == so["iters"] and rel(xg, xo) < 1e-9


def test_pcg_given_x0_and_stall(gpu_ctx_factory):
    _say("pcg-x0-stall")
    d = load_golden("geno_n150_odd")
    c = d["cohort"]
    g = device_for(d, gpu_ctx_factory)
    orc = golden_oracle(d)
    b = np.column_stack([c.X, d["y"]])
    x0 = np.random.default_rng(2).standard_normal(b.shape) * 0.01
    xg, _ = g.pcg(b, tol=1e-10, x0_mode=G.X0_GIVEN, x0=x0, rule=G.RULE_ALL_COLUMNS, stagnation=False)
    xo, _ = orc.pcg(b, tol=1e-10, x0_mode=2, x0=x0, rule=1, stagnation=False)
    assert rel(xg, xo) < 1e-8
    th = (1e-9, 10.0, 1e-4, 0.0, 1e-4)
    g.set_varcomp(th)
    ill = golden_oracle(d, theta=th)
    _, sg = g.pcg(b, tol=1e-14, max_iter=2000, rule=G.RULE_MAX_RATIO, stagnation=True)
    _, so = ill.pcg(b, tol=1e-14, max_iter=2000, rule=0, stagnation=True)
    assert sg["stalled"] == 1 and so["stalled"] == 1
    _, sg = g.pcg(b, tol=1e-14, max_iter=37, rule=G.RULE_MAX_RATIO, stagnation=False)
    assert sg["iters"] == 37


@pytest.mark.parametrize("name", GENO + ["grm_n180"])
def test_gls_blup_fp64_matches_oracle(name, gpu_ctx_factory):
    _say("gls", name)
    d = load_golden(name)
    g = device_for(d, gpu_ctx_factory)
    orc = golden_oracle(d)
    rg = g.gls_blup(d["y"], tol=1e-9, stagnation=False)
    ro = orc.gls_blup(d["y"], tol=1e-9, stagnation=False)
    for key in ("beta", "uhat", "bhat", "vinv_r"):
        assert rel(rg[key], ro[key]) < 1e-7, key
    # and the compiled reference's FP32 outputs to its own tolerance
    rg5 = g.gls_blup(d["y"], tol=1e-5)
    assert rel(rg5["beta"], d["ref_beta"]) < 5e-4
    assert rel(rg5["uhat"], d["ref_uhat"]) < 5e-4
    assert rel(rg5["bhat"], d["ref_bhat"]) < 5e-4


def test_mixed_solve_within_1e6_of_fp64(gpu_ctx_factory):
    _say("mixed-solve")
    d = load_golden("geno_qc_n257")
    g = device_for(d, gpu_ctx_factory)
    c = d["cohort"]
    rng = np.random.default_rng(7)
    b = np.column_stack([d["y"], rng.choice([-1.0, 1.0], size=(c.nrec, 31))])
    x64, s64 = g.pcg(b, tol=1e-9, rule=G.RULE_PER_COLUMN, precision=G.PREC_FP64)
    xm, sm = g.pcg(b, tol=1e-9, rule=G.RULE_PER_COLUMN, precision=G.PREC_MIXED)
    errs = np.linalg.norm(xm - x64, axis=0) / np.linalg.norm(x64, axis=0)
    assert errs.max() < 1e-6, errs.max()


@pytest.mark.parametrize("tol", [1e-6, 1e-9])
def test_mixed_relaxed_digits_within_1e6(tol, gpu_ctx_factory):
    """Mixed precision lowers the contraction's digits as the residual falls (inexact
    Krylov relaxation); the solution stays within 1e-6 of FP64, the true residual
    within a small multiple of tol, and fewer int8 operations are issued than with the
    fixed 4 digits (GLOBE_PCG_FIXED_DIGITS)."""
    _say("mixed-relaxed", tol)
    d = load_golden("geno_qc_n257")
    g = device_for(d, gpu_ctx_factory)
    c = d["cohort"]
    rng = np.random.default_rng(17)
    b = np.column_stack([d["y"], rng.choice([-1.0, 1.0], size=(c.nrec, 31))])
    x64, _ = g.pcg(b, tol=tol, rule=G.RULE_PER_COLUMN, precision=G.PREC_FP64)
    issued = {}
    sols = {}
    for name, flags in (("relaxed", 0), ("fixed", G.PCG_FIXED_DIGITS)):
        g.profile(True)
        sols[name], st = g.pcg(b, tol=tol, rule=G.RULE_PER_COLUMN, precision=G.PREC_MIXED, flags=flags)
        ks = [g.kernel_stats(0), g.kernel_stats(1)]
        issued[name] = sum(k["issued"] * k["launches"] for k in ks)
        g.profile(False)
        errs = np.linalg.norm(sols[name] - x64, axis=0) / np.linalg.norm(x64, axis=0)
        assert errs.max() < 1e-6, (name, errs.max())
        # true residual of the mixed solution, with the FP64 operator: within the PCG
        # tolerance plus the gap the quantised operator leaves (~1e-8 at 4 digits)
        res = g.computeVx_multi(sols[name]) - b
        assert (np.linalg.norm(res, axis=0) / np.linalg.norm(b, axis=0)).max() < 3 * tol + 3e-8, name
    # at tol 1e-6 the relaxation drops digits; at 1e-9 its bound keeps 4 digits until the
    # residual is below 5e-6, which the batched iterations of this small cohort may skip
    assert issued["relaxed"] < issued["fixed"] if tol == 1e-6 else issued["relaxed"] <= issued["fixed"]


@pytest.mark.parametrize("nmc", [7, 15])
def test_aireml_eval_matches_oracle(nmc, gpu_ctx_factory):
    """Gradient, trace and AI with injected probes; 15 = the reference's numMC default."""
    _say("aireml-eval", nmc)
    d = load_golden("grm_n180")
    th = (0.5, 0.5, 0.3, 0.05, 0.3)
    g = device_for(d, gpu_ctx_factory, theta=th)
    orc = golden_oracle(d, theta=th)
    c = d["cohort"]
    rng = np.random.default_rng(11)
    probes = rng.choice([-1.0, 1.0], size=(c.nrec, nmc))
    og = g.aireml_eval(d["y"], probes, probes, tol=1e-10)
    oo = orc.aireml_eval(d["y"], probes, probes, tol=1e-10)
    for key in ("first", "trace", "grad", "ai"):
        assert rel(og[key], oo[key]) < 1e-6, key
    # reference-style Gaussian tilde-V probes with x0 = weights = tilde-V^-1 probe
    z1, z2, z3 = rng.standard_normal(c.n), rng.standard_normal(2 * c.n), rng.standard_normal(c.nrec)
    yv = orc.sample_tildeV(z1, z2, z3)[:, None]
    wv = orc.tildeV_solve(yv)
    og = g.aireml_eval(d["y"], yv, wv, wv, tol=1e-10)
    oo = orc.aireml_eval(d["y"], yv, wv, wv, tol=1e-10)
    assert rel(og["grad"], oo["grad"]) < 1e-6 and rel(og["ai"], oo["ai"]) < 1e-6


@pytest.mark.parametrize("kind", ["rademacher", "gaussian"])
def test_device_probes_match_oracle(kind, gpu_ctx_factory):
    """Probes drawn on the device (Philox4x32-10) equal the oracle's restatement;
    Gaussian tilde-V weights equal tilde-V^-1 probe. (Exact unit probes are exercised by
    the reduced model's exact-trace fit, tests/test_reduced.py.)"""
    _say("device-probes", kind)
    d = load_golden("geno_n150_odd")
    th = (0.6, 0.4, 0.4, 0.2, 0.2)
    g = device_for(d, gpu_ctx_factory, theta=th)
    orc = golden_oracle(d, theta=th)
    c = d["cohort"]
    seed = int(g.L.globe_probe_seed(20261203, 3))
    nmc = 9
    Y, Wt = g.aireml_probes(kind, nmc, seed)
    want = O.probes({"gaussian": 0, "rademacher": 1}[kind], c.indiv, c.times, th, nmc, seed)
    if kind == "gaussian":
        assert rel(Y, want) < 1e-13
        assert rel(Wt, orc.tildeV_solve(Y)) < 1e-12
    else:
        assert np.array_equal(Y, want) and np.array_equal(Wt, Y)


def test_invalid_cohorts_rejected(gpu_ctx_factory):
    _say("invalid")
    g = gpu_ctx_factory()
    t = np.zeros(4)
    X = np.ones((4, 2))
    with pytest.raises(G.GlobeError, match="records of individual 1 are not contiguous"):
        g.set_cohort(np.array([0, 1, 0, 1]), t, X)  # individual 0 appears in two runs
    ind = np.array([0, 0, 1, 3], dtype=np.int32)  # id 3 outside [0, n = 3)
    Xf = np.asfortranarray(X)
    assert g.L.globe_set_cohort(g.h, 4, 3, 1, ind.ctypes.data, t.ctypes.data, Xf.ctypes.data, 4) == 1
    assert b"individual index 3 at record 3 outside [0, n)" in g.L.globe_last_error(g.h)
    ind = np.array([0, 0, 2, 2], dtype=np.int32)  # n = 3 but id 1 never occurs
    assert g.L.globe_set_cohort(g.h, 4, 3, 1, ind.ctypes.data, t.ctypes.data, Xf.ctypes.data, 4) == 1
    g.set_cohort(np.array([0, 0, 1, 1], dtype=np.int32), t, X)
    with pytest.raises(G.GlobeError):  # sig2e must be positive
        g.set_varcomp((0.0, 0.1, 0.1, 0.0, 0.1))
    with pytest.raises(G.GlobeError):  # no genetics yet
        g.set_varcomp((0.5, 0.1, 0.1, 0.0, 0.1))
        g.computeVx_multi(np.ones(4))


@pytest.mark.parametrize("big_block", [False, True])
def test_record_tiles_and_large_block_fallback(big_block, gpu_ctx_factory):
    """Cohorts spanning many record tiles (2048-record LDS tiles of whole individuals)
    and, with one individual holding more than 2048 records, the thread-per-individual
    kernels: operator, preconditioner and a PCG solve against the oracle."""
    import oracle as O
    _say("record-tiles", big_block)
    d = synth.make_geno_case(n=1400, m=700, lam=6, seed=77)
    c = d.cohort
    if big_block:
        counts = np.bincount(c.indiv, minlength=c.n)
        counts[5] = 2100
        indiv = np.repeat(np.arange(c.n, dtype=np.int32), counts)
        rng = np.random.default_rng(5)
        times = synth.f32(rng.uniform(0.0, 10.0, size=indiv.shape[0]))
        covars = np.stack([np.ones(indiv.shape[0]), (indiv % 2).astype(np.float64),
                           synth.f32(40.0 + (indiv % 30))], axis=1)
        c = synth.Cohort(n=c.n, indiv=indiv, times=times, covars=covars)
        d.cohort = c
        d.y = rng.standard_normal(c.nrec)
    g = gpu_ctx_factory()
    g.set_cohort(c.indiv, c.times, c.X)
    g.prepGenotypeInput(d.bed_rows, 0.01, 0.2)
    g.set_varcomp(d.theta)
    orc = O.oracle_from_synth(d)
    x = np.random.default_rng(9).standard_normal((c.nrec, 3))
    assert rel(g.computeVx_multi(x), orc.vx(x)) < 1e-12
    q = orc.precond()
    assert rel(g.applyPreconditioner(x), orc.precond_apply(q, x)) < 1e-11
    b = np.asarray(d.y)[:, None]
    xg, _ = g.pcg(b, tol=1e-9, rule=G.RULE_PER_COLUMN, precision=G.PREC_FP64, stagnation=False)
    xo, _ = orc.pcg(b, tol=1e-9, rule=2, stagnation=False, q=q)
    assert rel(xg, xo) < 1e-7


@pytest.mark.parametrize("k", [1, 3, 5, 8, 9, 16, 17, 20, 21, 24, 25, 32, 33, 52, 64])
def test_grm_product_widths(k, gpu_ctx_factory):
    """Tiled GRM product: one pass of 16 columns (1..16), of 16 plus one or two groups of
    4 on the 4 x 4 MFMA (17..20, 21..24), of two column groups sharing each tile (25..32),
    and several passes (33..64; 52 = 32 + 20)."""
    _say("grm-widths", k)
    d = load_golden("grm_n180")
    g = device_for(d, gpu_ctx_factory)
    orc = golden_oracle(d)
    x = np.random.default_rng(k).standard_normal((d["cohort"].nrec, k))
    assert rel(g.computeVx_multi(x), orc.vx(x)) < 1e-12
    assert rel(g.computeVx_multi(x, multGRM_uhat=True), orc.vx(x, grm_only=True)) < 1e-12


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 8])
def test_grm_few_column_product_multi_tile(k, gpu_ctx_factory):
    """Products of k <= 8 columns read each stored tile once (row strip in registers,
    transposed strip through per-tile partials, both on the 4 x 4 FP64 MFMA). n = 3,000:
    47 tile strips, so each of a strip's 8 chunks holds several tiles and both ring slots
    are used. Against sig2g G J^T x from the dense G, and against the same columns of a
    16-column product (the strip kernel)."""
    _say("grm-few-columns", k)
    n = 3000
    rng = np.random.default_rng(40 + k)
    c = synth.simulate_cohort(n, 0.5, rng)
    gcase = synth.make_reduced_case(n=n, m=500, seed=41)
    G = synth.grm_dense(gcase.grm_lt, n)
    theta = (0.6, 0.4, 0.4, 0.2, 0.2)
    g = gpu_ctx_factory()
    g.set_cohort(c.indiv, c.times, c.X)
    g.readFullGRM(gcase.grm_lt)
    g.set_varcomp(theta)
    x = rng.standard_normal((c.nrec, 16))
    J = np.zeros((c.nrec, n))
    J[np.arange(c.nrec), c.indiv] = 1.0
    want = theta[1] * G @ (J.T @ x[:, :k])
    got = g.computeVx_multi(np.asfortranarray(x[:, :k]), multGRM_uhat=True)
    assert rel(got, want) < 1e-12
    wide = g.computeVx_multi(np.asfortranarray(x), multGRM_uhat=True)[:, :k]
    assert rel(got, wide) < 1e-13


@pytest.mark.parametrize("k", [17, 20, 22, 24])
def test_grm_four_column_groups_multi_tile(k, gpu_ctx_factory):
    """17-24 columns: 16 on the 16 x 16 FP64 MFMA plus groups of 4 on the 4 x 4 one. n =
    3,000 (47 tile strips, tiles read both ways). Against sig2g G J^T x from the dense G,
    and against the same columns of a 32-column product (two 16-column groups)."""
    _say("grm-4col-groups", k)
    n = 3000
    rng = np.random.default_rng(60 + k)
    c = synth.simulate_cohort(n, 0.5, rng)
    gcase = synth.make_reduced_case(n=n, m=500, seed=61)
    G = synth.grm_dense(gcase.grm_lt, n)
    theta = (0.6, 0.4, 0.4, 0.2, 0.2)
    g = gpu_ctx_factory()
    g.set_cohort(c.indiv, c.times, c.X)
    g.readFullGRM(gcase.grm_lt)
    g.set_varcomp(theta)
    x = rng.standard_normal((c.nrec, 32))
    J = np.zeros((c.nrec, n))
    J[np.arange(c.nrec), c.indiv] = 1.0
    want = theta[1] * G @ (J.T @ x[:, :k])
    got = g.computeVx_multi(np.asfortranarray(x[:, :k]), multGRM_uhat=True)
    assert rel(got, want) < 1e-12
    wide = g.computeVx_multi(np.asfortranarray(x), multGRM_uhat=True)[:, :k]
    assert rel(got, wide) < 1e-13
    # each column on its own scale: a column of zeros and a tiny one stay exact
    x2 = x[:, :k].copy()
    x2[:, k - 1] = 0.0
    x2[:, 16] *= 1e-9
    got2 = g.computeVx_multi(np.asfortranarray(x2), multGRM_uhat=True)
    want2 = theta[1] * G @ (J.T @ x2)
    assert np.all(got2[:, k - 1] == 0.0)
    assert rel(got2[:, 16], want2[:, 16]) < 1e-12


def test_grm_builder_at_40k_rows_offsets(gpu_ctx_factory):
    """The builder forms only rows i >= j of each tile column, from a row offset (a
    multiple of 512) on; fewer rows let the contraction take more K-splits than the
    full cohort, and its partial-sum buffer must be sized for them (N = 40,000 and
    M = 80,000 SNPs, the AI-REML benchmark's shape, is the first size where they
    differ). The built GRM's products and diagonal match the genotype path."""
    _say("grm-builder-40k")
    rng = np.random.default_rng(31)
    c = synth.simulate_cohort(40_000, 0.5, rng)
    g = gpu_ctx_factory()
    g.set_cohort(c.indiv, c.times, c.X)
    g.synth_bed(80_000, seed=31, f_lo=0.05, f_hi=0.5, miss=0.0, min_maf=0.0, max_miss=1.0)
    g.set_varcomp((0.6, 0.4, 0.4, 0.2, 0.2))
    x = rng.standard_normal((c.nrec, 3))
    v_geno = g.computeVx_multi(x)
    d_geno = g.computeGeneticDiagonal()
    g.grm_from_geno()
    assert rel(g.computeVx_multi(x), v_geno) < 1e-6
    assert rel(g.computeGeneticDiagonal(), d_geno) < 1e-6


def test_grm_builder_matches_wwt(gpu_ctx_factory):
    """G = W W^T / M built on the device from the resident genotypes equals the numpy
    product to FP32 rounding, and the context then solves from it like a loaded GRM."""
    import oracle as O
    _say("grm-builder")
    d = synth.make_geno_case(n=333, m=700, lam=3, seed=21, miss=0.01)
    c = d.cohort
    g = gpu_ctx_factory()
    g.set_cohort(c.indiv, c.times, c.X)
    g.prepGenotypeInput(d.bed_rows, 0.01, 0.2)
    g.set_varcomp(d.theta)
    x = np.random.default_rng(4).standard_normal((c.nrec, 3))
    v_geno = g.computeVx_multi(x)
    g.grm_from_geno()
    W = synth.standardized(d.bed_rows[d.keep], c.n, d.freq)
    G = (W @ W.T) / W.shape[1]
    lt = g.grm_lt()
    ref = G[np.tril_indices(c.n)].astype(np.float32)
    assert np.max(np.abs(lt - ref)) <= 2.0 * np.finfo(np.float32).eps * np.max(np.abs(G))
    assert rel(g.computeVx_multi(x), v_geno) < 1e-6
    orc = O.Oracle(c.indiv, c.times, c.X, d.theta, grm_lt=lt, n=c.n)
    assert rel(g.computeVx_multi(x), orc.vx(x)) < 1e-12
    assert rel(g.computeGeneticDiagonal(), np.diag(G)) < 1e-6


@pytest.mark.parametrize("kind", ["geno", "grm"])
def test_cohort_of_another_size_releases_genetics(kind, gpu_ctx_factory):
    """Genotypes and GRMs are laid out for the N they were staged with: a cohort of
    another N releases them, so the next product fails with GLOBE_E_STATE (before any
    kernel reads buffers of the old size); a cohort of the same N keeps them."""
    _say("cohort-resize", kind)
    d = synth.make_geno_case(n=200, m=300, lam=2, seed=5) if kind == "geno" else \
        synth.make_grm_case(n=200, m=300, lam=2, seed=5)
    c = d.cohort
    g = gpu_ctx_factory()
    g.set_cohort(c.indiv, c.times, c.X)
    if kind == "geno":
        g.prepGenotypeInput(d.bed_rows, 0.01, 0.2)
    else:
        g.readFullGRM(d.grm_lt)
    g.set_varcomp(d.theta)
    x = np.random.default_rng(1).standard_normal((c.nrec, 2))
    y0 = g.computeVx_multi(x)
    # same N, records reordered within the cohort's rules: genetics stay
    g.set_cohort(c.indiv, c.times, c.X)
    assert np.array_equal(g.computeVx_multi(x), y0)
    for n_new in (150, 260):
        c2 = synth.simulate_cohort(n_new, 2.0, np.random.default_rng(n_new))
        g.set_cohort(c2.indiv, c2.times, c2.X)
        with pytest.raises(G.GlobeError) as e:
            g.computeVx_multi(np.ones((c2.nrec, 1)))
        assert e.value.code == 2  # GLOBE_E_STATE
    g.close()
