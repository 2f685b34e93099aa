"""globe_vx_dev_noalloc: the operator with every workspace in the static device scratch.

Through the C ABI, against globe_vx_dev on the same device buffers: bit-identical
results (the same kernels; every contraction candidate sums the same integers), on two
calls in a row with the second one smaller (fewer columns, then a smaller cohort in a
new context), so leftovers of an earlier call in the shared scratch cannot leak into a
later one; the packed-genotype and GRM paths in FP64, mixed precision and the FP64
vector-ALU path; a call larger than the scratch is refused with GLOBE_E_INVALID, and a
mixed GRM product whose digit tiles would have to be allocated with GLOBE_E_STATE.
"""
import numpy as np
import pytest

import globe_amd as G
from globe_amd import synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _vx_pair(g, x, mode, prec):
    """(globe_vx_dev, globe_vx_dev_noalloc) on the same device input."""
    rows = g.n if mode == G.VX_GRM_ONLY else g.nrec
    k = x.shape[1]
    xb = g.alloc(8 * g.nrec * k)
    xb.upload(np.asfortranarray(x))
    ya, yb = g.alloc(8 * rows * k), g.alloc(8 * rows * k)
    g.vx_dev(xb, ya, k, mode=mode, precision=prec)
    g._check(g.L.globe_vx_dev_noalloc(g.h, xb.ptr, g.nrec, k, yb.ptr, rows, mode, prec))
    g.synchronize()
    a = ya.download(np.zeros((rows, k), order="F"))
    b = yb.download(np.zeros((rows, k), order="F"))
    for buf in (xb, ya, yb):
        buf.free()
    return a, b


def _geno_ctx(factory, n, m, lam, seed):
    c = synth.simulate_cohort(n, lam, np.random.default_rng(seed))
    g = factory()
    g.set_cohort(c.indiv, c.times, c.X)
    g.synth_bed(m, seed, 0.01, 0.5, 0.01, 0.01, 0.05)
    g.set_varcomp((0.6, 0.4, 0.4, 0.2, 0.2))
    return g, c


def test_noalloc_matches_vx_dev_geno(gpu_ctx_factory):
    rng = np.random.default_rng(3)
    g, c = _geno_ctx(gpu_ctx_factory, 20000, 40000, 2.0, 5)
    # two calls in a row on one context, the second with fewer columns
    for prec, ks in ((G.PREC_MIXED, (32, 5)), (G.PREC_FP64, (9, 1)), (G.PREC_FP64_VALU, (4, 2))):
        for mode in (G.VX_FULL, G.VX_GRM_ONLY):
            for k in ks:
                _say("noalloc geno", prec, mode, k)
                x = rng.standard_normal((c.nrec, k))
                a, b = _vx_pair(g, x, mode, prec)
                assert np.array_equal(a, b), (prec, mode, k, float(np.abs(a - b).max()))
    g.close()
    # N = 200k from 40k SNPs: 8 columns need more than the scratch (refused, nothing is
    # written to Y), one column fits and still matches after the refusal
    g, c = _geno_ctx(gpu_ctx_factory, 200000, 40000, 1.0, 8)
    k = 8
    xb, yb = g.alloc(8 * c.nrec * k), g.alloc(8 * c.nrec * k)
    yb.upload(np.full((c.nrec, k), 7.0, order="F"))
    rc = g.L.globe_vx_dev_noalloc(g.h, xb.ptr, c.nrec, k, yb.ptr, c.nrec, G.VX_FULL, G.PREC_FP64)
    assert rc == 1 and b"scratch" in g.L.globe_last_error(g.h)
    g.synchronize()
    assert (yb.download(np.zeros((c.nrec, k), order="F")) == 7.0).all()
    xb.free()
    yb.free()
    for prec in (G.PREC_FP64, G.PREC_MIXED):
        _say("noalloc geno 200k", prec)
        a, b = _vx_pair(g, rng.standard_normal((c.nrec, 1)), G.VX_FULL, prec)
        assert np.array_equal(a, b)
    g.close()
    # a smaller cohort in a new context: the scratch holds the last call's data
    g, c = _geno_ctx(gpu_ctx_factory, 3000, 5000, 1.5, 6)
    for prec in (G.PREC_FP64, G.PREC_MIXED):
        _say("noalloc geno small", prec)
        x = rng.standard_normal((c.nrec, 3))
        a, b = _vx_pair(g, x, G.VX_FULL, prec)
        assert np.array_equal(a, b)
    g.close()


def test_noalloc_matches_vx_dev_grm(gpu_ctx_factory):
    d = synth.make_grm_case(n=700, m=900, lam=3, seed=9)
    c = d.cohort
    g = gpu_ctx_factory()
    g.set_cohort(c.indiv, c.times, c.X)
    g.readFullGRM(d.grm_lt)
    g.set_varcomp(d.theta)
    rng = np.random.default_rng(1)
    # the mixed product's digit tiles are built by the first (allocating) mixed product
    xb, yb = g.alloc(8 * c.nrec), g.alloc(8 * c.nrec)
    rc = g.L.globe_vx_dev_noalloc(g.h, xb.ptr, c.nrec, 1, yb.ptr, c.nrec, G.VX_FULL, G.PREC_MIXED)
    assert rc == 2
    g.synchronize()
    for prec, ks in ((G.PREC_FP64, (20, 3)), (G.PREC_MIXED, (33, 1))):
        for k in ks:
            _say("noalloc grm", prec, k)
            x = rng.standard_normal((c.nrec, k))
            a, b = _vx_pair(g, x, G.VX_FULL, prec)
            assert np.array_equal(a, b), (prec, k)
    g.close()
