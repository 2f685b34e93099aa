"""Pins the FP64 oracle (oracle/globe_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign or
index, a transposed operand, a wrong phantom mask) fails it:
  * PLINK decode / counts / QC: brute force over 2-bit codes in numpy (including
    N % 4 != 0 phantom pairs) and the kept-SNP counts printed by the compiled
    reference runPCGsolver (golden fixtures);
  * operator: an explicit dense V = sig2g J G J^T + blockdiag(U B U^T) + sig2e I with
    G = W W^T / M in numpy, the record-pair form, and symmetry;
  * preconditioner: numpy dense block inverse;
  * PCG / GLS / BLUP: numpy dense solve and GLS / Henderson identities, and the
    _beta/_uhat/_bhat.bin files written by the compiled reference (FP32, tol 1e-5);
  * AIREML terms: exact dense P, f = dV P y, the Hutchinson estimator with the
    same probes, AI = f^T P f / 2;
  * the tilde-V sampling formula: its exact covariance.
"""
import numpy as np
import pytest

import oracle as O
from globe_amd import synth
from helpers import dense_V, golden_oracle, load_golden, rel

GOLDEN_GENO = ["geno_n203", "geno_n150_odd", "geno_qc_n257"]


def brute_counts(rows, n):
    alt = np.zeros(rows.shape[0], dtype=np.int64)
    miss = np.zeros(rows.shape[0], dtype=np.int64)
    for s in range(rows.shape[0]):
        for i in range(n):
            code = (int(rows[s, i // 4]) >> (2 * (i % 4))) & 3
            if code == 0b00:
                alt[s] += 2
            elif code == 0b10:
                alt[s] += 1
            elif code == 0b01:
                miss[s] += 1
    return alt, miss


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 13, 64, 67])
def test_decode_and_counts_brute_force(n):
    rng = np.random.default_rng(n)
    rows = rng.integers(0, 256, size=(37, (n + 3) // 4), dtype=np.uint8)
    alt, miss = O.bed_counts(rows, n)
    ba, bm = brute_counts(rows, n)
    assert np.array_equal(alt, ba) and np.array_equal(miss, bm)
    assert [O.lib().orc_decode(c) for c in range(4)] == [2, -1, 1, 0]


def test_qc_decision_matches_float32_rule():
    rng = np.random.default_rng(3)
    n = 97
    dos = synth.simulate_dosages(n, 400, rng, f_lo=0.0, f_hi=0.1, miss=0.05)
    rows = synth.pack_bed(dos)
    keep, maf, mr, kept = O.bed_qc(rows, n, 0.03, 0.06)
    k2, maf2, mr2, _, _ = synth.qc_reference(rows, n, 0.03, 0.06)
    assert np.array_equal(keep, k2) and np.array_equal(maf, maf2) and np.array_equal(mr, mr2)
    assert 0 < kept < 400


@pytest.mark.parametrize("name", GOLDEN_GENO)
def test_qc_kept_count_matches_reference(name):
    d = load_golden(name)
    keep, _, _, kept = O.bed_qc(d["bed_rows"], d["cohort"].n, float(d["min_maf"]), float(d["max_miss"]))
    assert kept == int(d["ref_kept"])


@pytest.mark.parametrize("name", ["geno_n203", "grm_n180"])
def test_vx_matches_explicit_dense_V(name):
    d = load_golden(name)
    orc = golden_oracle(d)
    c = d["cohort"]
    if "grm_lt" in d:
        G = synth.grm_dense(d["grm_lt"], c.n)
    else:
        W = synth.standardized(d["bed_rows"][d["keep"]], c.n, d["freq"])
        G = W @ W.T / W.shape[1]
        assert np.allclose(orc.genetic_diag(), np.diag(G), rtol=1e-13, atol=1e-13)
    V, J = dense_V(c, G, tuple(d["theta"]))
    x = np.random.default_rng(1).standard_normal((c.nrec, 3))
    assert rel(orc.vx(x), V @ x) < 1e-13
    assert rel(orc.vx_dense(x), V @ x) < 1e-13
    sg = d["theta"][1]
    assert rel(orc.vx(x, grm_only=True), sg * G @ J.T @ x) < 1e-13
    # symmetry of the operator
    y = np.random.default_rng(2).standard_normal((c.nrec, 1))
    assert abs((y.T @ orc.vx(x[:, :1]))[0, 0] - (x[:, :1].T @ orc.vx(y))[0, 0]) < 1e-10
    # FP32 restatement kept for CPU timing agrees to FP32 rounding
    assert rel(orc.vx_ref32(x), V @ x) < 1e-5
    if "grm_lt" in d:
        # and so does the record-pair form of runAIREML's operator (computeVx_multi_VC),
        # every pair V_ij built on the fly; dropping the time term or the residual fails
        y32 = orc.vx_vc_ref32(x)
        assert rel(y32, V @ x) < 1e-5
        se = d["theta"][0]
        assert rel(y32, (V - se * np.eye(c.nrec)) @ x) > 1e-3


def test_geno_Wa_matches_dense_standardized_genotypes():
    """The phenotype recipe's W a (standardized genotypes of the kept SNPs times a
    SNP-space vector) against the dense FP64 W built independently by synth.standardized,
    with missing calls and QC-dropped SNPs; a W without its centering fails."""
    d = synth.make_geno_case(n=301, m=700, lam=3, seed=41, miss=0.02)
    n = d.cohort.n
    W = synth.standardized(d.bed_rows[d.keep], n, d.freq)
    a = np.random.default_rng(6).standard_normal(W.shape[1])
    got = O.oracle_from_synth(d).geno_Wa(a)
    assert rel(got, W @ a) < 1e-12
    codes = synth.unpack_codes(d.bed_rows[d.keep], n)
    uncentred = synth.DOSAGE_OF_CODE[codes].astype(np.float64).T
    assert rel(got, uncentred @ a) > 1e-3


def test_preconditioner_is_dense_block_inverse():
    d = load_golden("geno_n150_odd")
    orc = golden_oracle(d)
    c = d["cohort"]
    se, sg, sb0, sb01, sb1 = d["theta"]
    gd = orc.genetic_diag()
    t = c.times
    same = c.indiv[:, None] == c.indiv[None, :]
    Q = same * (sg * gd[c.indiv][:, None] + sb0 + sb01 * (t[:, None] + t[None, :]) + sb1 * np.outer(t, t))
    Q += se * np.eye(c.nrec)
    r = np.random.default_rng(5).standard_normal((c.nrec, 2))
    z = orc.precond_apply(orc.precond(), r)
    assert rel(Q @ z, r) < 1e-12


@pytest.mark.parametrize("rule", [0, 1, 2])
def test_pcg_matches_dense_solve(rule):
    d = load_golden("geno_n203")
    orc = golden_oracle(d)
    c = d["cohort"]
    W = synth.standardized(d["bed_rows"][d["keep"]], c.n, d["freq"])
    V, _ = dense_V(c, W @ W.T / W.shape[1], tuple(d["theta"]))
    b = np.column_stack([c.X, d["y"]])
    x, st = orc.pcg(b, tol=1e-11, max_iter=500, rule=rule, stagnation=False)
    xs = np.linalg.solve(V, b)
    assert rel(x, xs) < 1e-9
    assert st["iters"] < 100
    # a looser tolerance stops earlier and each column's residual respects it
    x2, st2 = orc.pcg(b, tol=1e-4, max_iter=500, rule=rule, stagnation=False)
    res = np.linalg.norm(V @ x2 - b, axis=0) / np.linalg.norm(b, axis=0)
    assert st2["iters"] < st["iters"] and res.max() < 1e-4


def test_pcg_x0_modes_and_stall_exit():
    d = load_golden("geno_n150_odd")
    orc = golden_oracle(d)
    c = d["cohort"]
    b = np.column_stack([c.X, d["y"]])
    x1, _ = orc.pcg(b, tol=1e-10, x0_mode=1, rule=1, stagnation=False)
    x0 = x1 + 1e-3 * np.random.default_rng(0).standard_normal(x1.shape)
    x2, st = orc.pcg(b, tol=1e-10, x0_mode=2, x0=x0, rule=1, stagnation=False)
    assert rel(x2, x1) < 1e-8
    # an ill-conditioned V (sig2e 1e-9 against sig2g 10 with G of rank < N) makes PCG
    # stall; the 10-iteration / 5e-3 rule ends it (PCGsolver.cpp:1402-1415)
    ill = golden_oracle(d, theta=(1e-9, 10.0, 1e-4, 0.0, 1e-4))
    _, st3 = ill.pcg(b, tol=1e-14, max_iter=2000, rule=0, stagnation=True)
    assert st3["stalled"] == 1 and st3["iters"] < 2000
    _, st4 = ill.pcg(b, tol=1e-14, max_iter=150, rule=0, stagnation=False)
    assert st4["stalled"] == 0 and st4["iters"] == 150


def test_gls_blup_closed_forms():
    d = load_golden("grm_n180")
    orc = golden_oracle(d)
    c = d["cohort"]
    G = synth.grm_dense(d["grm_lt"], c.n)
    th = tuple(d["theta"])
    V, J = dense_V(c, G, th)
    Vi = np.linalg.inv(V)
    X, y = c.X, d["y"]
    beta = np.linalg.solve(X.T @ Vi @ X, X.T @ Vi @ y)
    vr = Vi @ (y - X @ beta)
    r = orc.gls_blup(y, tol=1e-12, stagnation=False)
    assert rel(r["beta"], beta) < 1e-9
    assert rel(r["uhat"], th[1] * G @ J.T @ vr) < 1e-8
    bh = np.zeros(2 * c.n)
    np.add.at(bh, 2 * c.indiv, vr * (th[2] + th[3] * c.times))
    np.add.at(bh, 2 * c.indiv + 1, vr * (th[3] + th[4] * c.times))
    assert rel(r["bhat"], bh) < 1e-8
    # Henderson: u-hat and b-hat are the BLUPs E[u | y] = sig2g G J^T V^-1 (y - X beta)


@pytest.mark.parametrize("name", GOLDEN_GENO + ["grm_n180"])
def test_oracle_matches_compiled_reference_outputs(name):
    """The reference is FP32 with PCG tol 1e-5: agreement at ~1e-4 is the bar."""
    d = load_golden(name)
    orc = golden_oracle(d)
    r = orc.gls_blup(d["y"], tol=1e-5, stagnation=True)
    assert rel(r["beta"], d["ref_beta"]) < 5e-4
    assert rel(r["uhat"], d["ref_uhat"]) < 5e-4
    assert rel(r["bhat"], d["ref_bhat"]) < 5e-4
    # the oracle's FP64 PCG, same control flow, stops within one iteration of the reference's
    assert abs(r["iters"] - int(d["ref_iters"][0])) <= 2


def test_aireml_terms_match_dense_P():
    d = load_golden("grm_n180")
    c = d["cohort"]
    th = (0.5, 0.5, 0.3, 0.05, 0.3)
    orc = golden_oracle(d, theta=th)
    G = synth.grm_dense(d["grm_lt"], c.n)
    V, J = dense_V(c, G, th)
    Vi = np.linalg.inv(V)
    X = c.X
    P = Vi - Vi @ X @ np.linalg.solve(X.T @ Vi @ X, X.T @ Vi)
    t = c.times
    same = (c.indiv[:, None] == c.indiv[None, :]).astype(float)
    dV = [np.eye(c.nrec), J @ G @ J.T, same, same * (t[:, None] + t[None, :]), same * np.outer(t, t)]
    y = d["y"]
    Py = P @ y
    f = np.column_stack([D @ Py for D in dV])
    rng = np.random.default_rng(9)
    nmc = 6
    probes = rng.choice([-1.0, 1.0], size=(c.nrec, nmc))
    out = orc.aireml_eval(y, probes, probes, tol=1e-12)
    assert rel(out["first"], Py @ f) < 1e-8
    tr = np.array([np.mean([(D @ P @ probes[:, s]) @ probes[:, s] for s in range(nmc)]) for D in dV])
    assert rel(out["trace"], tr) < 1e-8
    assert rel(out["grad"], 0.5 * (Py @ f - tr)) < 1e-8
    assert rel(out["ai"], 0.5 * f.T @ P @ f) < 1e-8
    # the Hutchinson estimate converges to tr(P dV) (loose, statistical)
    exact = np.array([np.trace(P @ D) for D in dV])
    many = rng.choice([-1.0, 1.0], size=(c.nrec, 400))
    est = np.array([np.mean(np.einsum("ij,ij->j", many, D @ P @ many)) for D in dV])
    assert rel(est, exact) < 0.1


def test_tildeV_sampling_covariance_and_solve():
    d = load_golden("geno_n150_odd")
    c = d["cohort"]
    th = (0.6, 0.4, 0.4, 0.2, 0.2)
    orc = golden_oracle(d, theta=th)
    se, sg, sb0, sb01, sb1 = th
    t = c.times
    same = (c.indiv[:, None] == c.indiv[None, :]).astype(float)
    Vt = same * (sg + sb0 + sb01 * (t[:, None] + t[None, :]) + sb1 * np.outer(t, t)) + se * np.eye(c.nrec)
    # the sampling formula is linear in (z1, z2, z3): its covariance is A A^T
    nz = c.n + 2 * c.n + c.nrec
    A = np.zeros((c.nrec, nz))
    for q in range(nz):
        e = np.zeros(nz)
        e[q] = 1.0
        A[:, q] = orc.sample_tildeV(e[: c.n], e[c.n: 3 * c.n], e[3 * c.n:])
    assert rel(A @ A.T, Vt) < 1e-12
    yv = np.random.default_rng(4).standard_normal((c.nrec, 2))
    assert rel(orc.tildeV_solve(yv), np.linalg.solve(Vt, yv)) < 1e-12


def test_philox_known_answers_and_probe_moments():
    """Philox4x32-10 against the published known-answer vectors (Random123), then the
    probe generator's moments: Rademacher signs balanced, Gaussian tilde-V probes with
    covariance tilde-V (statistical, loose)."""
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, want in kat:
        got = O.philox4x32_10(*ctr, *key)
        assert [int(x) for x in got] == list(want)
    c = synth.simulate_cohort(12, 2.0, np.random.default_rng(5))
    th = (0.6, 0.4, 0.4, 0.2, 0.2)
    z = O.probes(1, c.indiv, c.times, th, 64, 12345)
    assert set(np.unique(z)) == {-1.0, 1.0} and abs(z.mean()) < 0.05
    yv = O.probes(0, c.indiv, c.times, th, 4000, 777)
    se, sg, sb0, sb01, sb1 = th
    t = c.times
    same = (c.indiv[:, None] == c.indiv[None, :]).astype(float)
    Vt = same * (sg + sb0 + sb01 * (t[:, None] + t[None, :]) + sb1 * np.outer(t, t)) + se * np.eye(c.nrec)
    emp = yv @ yv.T / yv.shape[1]
    assert np.abs(emp - Vt).max() < 0.1 * np.abs(Vt).max()
    e = O.probes(2, c.indiv, c.times, th, c.nrec, 1)
    assert np.allclose(e, np.sqrt(c.nrec) * np.eye(c.nrec))
