"""Mixed-precision dense GRM product (int8 digits of G and S0 on the matrix cores).

CPU tests pin the oracle's restatement (oracle.grm_mixed_product) against mathematics:
the balanced digit split is an identity on its range, the product with L digits
approaches the exact FP64 product by about 2^8 per digit and stays within the
quantisation bound, and L = 4 agrees with FP64 to FP32-rounding scale. GPU tests
compare the device product with the restatement bit for bit (one record per
individual, so S0 = x exactly), and the mixed PCG and AI-REML with the FP64 ones within
the stated tolerances (solutions 1e-6, variance components 1e-4).
"""
import numpy as np
import pytest

import globe_amd as G
import oracle as O
from globe_amd import synth
from helpers import rel


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def case():
    # 300 individuals: 5 tile rows, chunks of 1 and 2 tile columns
    return synth.make_reduced_case(n=300, m=900, seed=5, theta=(0.5, 0.5))


# ------------------------------------------------------------------------------ CPU
def test_digit_split_is_exact():
    rng = np.random.default_rng(1)
    v = np.concatenate([rng.integers(-(1 << 30), (1 << 30) + 1, size=5000),
                        [0, 1, -1, 127, 128, -128, -129, (1 << 30), -(1 << 30), (1 << 24) - 1]])
    d = O.digits4(v)
    back = sum(d[a] * (256 ** (3 - a)) for a in range(4))
    assert np.array_equal(back, v)
    assert np.abs(d[0]).max() <= 64
    assert d[1:].min() >= -128 and d[1:].max() <= 127


def test_mixed_product_converges_to_fp64(case):
    d = case
    n = d.n
    Gd = synth.grm_dense(d.grm_lt, n)
    x = np.random.default_rng(2).standard_normal((n, 5))
    exact = Gd @ x
    errs = []
    for L in (2, 3, 4):
        got = O.grm_mixed_product(d.grm_lt, n, x, L)
        err = np.abs(got - exact).max()
        # quantisation bound: each operand truncated to L digits is off by at most
        # 2^(1 - 8L) of its scale (row max off the diagonal, column max of x)
        off = np.abs(Gd - np.diag(np.diag(Gd))).max(axis=1)
        bound = 2.0 ** (2 - 8 * L) * (np.abs(Gd).sum(axis=1) * np.abs(x).max() + n * off.max() * np.abs(x).max())
        assert err <= bound.max(), (L, err, bound.max())
        errs.append(err)
    # each digit buys about 2^8
    assert errs[1] < errs[0] / 64 and errs[2] < errs[1] / 64
    assert rel(O.grm_mixed_product(d.grm_lt, n, x, 4), exact) < 1e-8
    # the diagonal goes through FP64: a G that is diagonal only is exact
    lt_diag = np.zeros_like(d.grm_lt)
    idx = np.arange(n)
    lt_diag[idx * (idx + 1) // 2 + idx] = d.grm_lt[idx * (idx + 1) // 2 + idx]
    got = O.grm_mixed_product(lt_diag, n, x, 2)
    assert np.array_equal(got, np.diag(synth.grm_dense(lt_diag, n))[:, None] * x)


# ------------------------------------------------------------------------------ GPU
def _device(d, gpu_ctx_factory, theta2=(0.5, 0.5)):
    g = gpu_ctx_factory()
    n = d.n
    g.set_cohort(np.arange(n, dtype=np.int32), np.zeros(n), d.covars)
    g.readFullGRM(d.grm_lt)
    g.set_varcomp((theta2[0], theta2[1], 0.0, 0.0, 0.0))
    return g


@pytest.mark.gpu
@pytest.mark.parametrize("k", [1, 16, 21, 33])
def test_mixed_grm_product_bit_exact(k, case, gpu_ctx_factory):
    _say("grm-mixed-exact", k)
    d = case
    g = _device(d, gpu_ctx_factory, (0.5, 0.7))
    x = np.random.default_rng(10 + k).standard_normal((d.n, k))
    x[:, 0] *= 1e-3  # columns of different scales
    for L in (4, 3, 2):
        g.set_mixed_digits(L)
        got = g.computeVx_multi(x, multGRM_uhat=True, precision=G.PREC_MIXED)
        want = O.grm_mixed_product(d.grm_lt, d.n, x, L, sig2g=0.7)
        assert np.array_equal(got, want), (L, np.abs(got - want).max())
    g.set_mixed_digits(4)
    mixed = g.computeVx_multi(x, multGRM_uhat=True, precision=G.PREC_MIXED)
    fp64 = g.computeVx_multi(x, multGRM_uhat=True, precision=G.PREC_FP64)
    assert rel(mixed, fp64) < 1e-8


@pytest.mark.gpu
def test_mixed_grm_small_and_zero_columns(gpu_ctx_factory):
    """n = 70 (two tile rows, empty chunks) and an all-zero column."""
    _say("grm-mixed-small")
    d = synth.make_reduced_case(n=70, m=300, seed=8)
    g = _device(d, gpu_ctx_factory)
    x = np.random.default_rng(3).standard_normal((d.n, 3))
    x[:, 1] = 0.0
    got = g.computeVx_multi(x, multGRM_uhat=True, precision=G.PREC_MIXED)
    want = O.grm_mixed_product(d.grm_lt, d.n, x, 4, sig2g=0.5)
    assert np.array_equal(got, want)
    assert np.all(got[:, 1] == 0.0)


@pytest.mark.gpu
def test_mixed_grm_pcg_within_1e6(case, gpu_ctx_factory):
    _say("grm-mixed-pcg")
    d = case
    g = _device(d, gpu_ctx_factory, (0.6, 0.4))
    b = np.random.default_rng(4).standard_normal((d.n, 8))
    x64, _ = g.pcg(b, tol=1e-9, rule=G.RULE_PER_COLUMN, precision=G.PREC_FP64, stagnation=False)
    xm, st = g.pcg(b, tol=1e-9, rule=G.RULE_PER_COLUMN, precision=G.PREC_MIXED, stagnation=False)
    assert rel(xm, x64) < 1e-6
    res = g.computeVx_multi(xm) - b
    assert (np.linalg.norm(res, axis=0) / np.linalg.norm(b, axis=0)).max() < 1e-7


@pytest.mark.gpu
def test_mixed_grm_aireml_within_1e4(case, gpu_ctx_factory):
    """Reduced-model AI-REML from the same device probes: mixed vs FP64 variance
    components."""
    _say("grm-mixed-aireml")
    d = case
    out = {}
    for prec in (G.PREC_FP64, G.PREC_MIXED):
        g = _device(d, gpu_ctx_factory)
        out[prec] = g.aireml_fit(d.y, (0.5, 0.5), max_iter=8, tol=1e-3, num_mc=15, probes="rademacher",
                                 seed=77, pcg_tol=1e-8, precision=prec, reduced=True)
    t64, tm = np.asarray(out[G.PREC_FP64]["theta"]), np.asarray(out[G.PREC_MIXED]["theta"])
    assert np.max(np.abs(tm - t64) / np.abs(t64)) < 1e-4, (t64, tm)


@pytest.mark.gpu
def test_grm_products_at_70k_individuals(gpu_ctx_factory):
    """N = 70,000 (1,094 tile rows): the mixed product sums its int32 digit classes over
    5 chunks of at most 256 tile columns (the int32 bound), the FP64 product over its 4;
    both agree, for 1 to 40 columns (two passes), with 4 digits within 1e-8, and the
    read-once 5-column FP64 product agrees with the strip kernel. The GRM is built on the
    device from 2,000 synthetic SNPs (9.6 GB of FP32 tiles, 19 GB of digit tiles)."""
    _say("grm-70k")
    rng = np.random.default_rng(70)
    c = synth.simulate_cohort(70_000, 0.2, rng)
    g = gpu_ctx_factory()
    g.set_cohort(c.indiv, c.times, c.X)
    g.synth_bed(2_000, seed=70, f_lo=0.05, f_hi=0.5, miss=0.0, min_maf=0.0, max_miss=1.0)
    g.set_varcomp((0.6, 0.4, 0.4, 0.2, 0.2))
    g.grm_from_geno()
    for k in (1, 40):
        x = rng.standard_normal((c.nrec, k))
        v64 = g.computeVx_multi(x, multGRM_uhat=True, precision=G.PREC_FP64)
        vm = g.computeVx_multi(x, multGRM_uhat=True, precision=G.PREC_MIXED)
        assert rel(vm, v64) < 1e-8, k
    # the read-once 4 x 4 MFMA product (5 columns) against the same columns of the
    # 16-column strip product, at 1,094 tile strips
    x = np.asfortranarray(rng.standard_normal((c.nrec, 16)))
    wide = g.computeVx_multi(x, multGRM_uhat=True, precision=G.PREC_FP64)[:, :5]
    five = g.computeVx_multi(np.asfortranarray(x[:, :5]), multGRM_uhat=True, precision=G.PREC_FP64)
    assert rel(five, wide) < 1e-13
