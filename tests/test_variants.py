"""Every candidate variant of the packed-genotype contraction, forced one at a time.

The tuner keeps the fastest candidate of each shape, so the variant that runs depends
on timing; these tests take that choice away (globe_force_variant) and run each
candidate of both passes: block shapes, split-K up to 64 ways with the XCD-aware
walk (S > 8), multi-piece stream-K grids, and every digit-column width (thin, and 1 to 4
wide fragments). All candidates sum the same integers, so V x must be bit-identical
across them; against the FP64 vector-ALU path (GLOBE_PREC_FP64_VALU, a different
algorithm: FP64 FMAs over the dosages) the 7-digit FP64 product is within 1e-12 and the
4-digit mixed product within its 30-bit quantisation (1e-7). The candidate sets at the
N = 20k / M = 40k cohort contain S > 8 and multi-piece stream-K (tests/test_plan_replay.py
checks that on the CPU, together with every candidate's bounds); the last test repeats
the comparison at the full N = 200k / M = 500k scale shape.
"""
import numpy as np
import pytest

import globe_amd as G
from globe_amd import synth
from helpers import rel

pytestmark = pytest.mark.gpu

# (precision, k): the digit-column widths 7k (FP64) and 4k (mixed) cover thin and 1-4 wide
# fragments, and FP64 k = 32 splits into two contractions of 18 and 14 columns
CASES = [(G.PREC_FP64, 1), (G.PREC_FP64, 4), (G.PREC_FP64, 9), (G.PREC_FP64, 13), (G.PREC_FP64, 32),
         (G.PREC_MIXED, 1), (G.PREC_MIXED, 8), (G.PREC_MIXED, 16), (G.PREC_MIXED, 24), (G.PREC_MIXED, 32)]
BOUND = {G.PREC_FP64: 1e-12, G.PREC_MIXED: 1e-7}


def _say(*a):
    print("[case]", *a, flush=True)


def _cohort_ctx(factory, n, m, lam, seed):
    rng = np.random.default_rng(seed)
    c = synth.simulate_cohort(n, lam, rng)
    g = factory()
    g.set_cohort(c.indiv, c.times, c.X)
    g.synth_bed(m, seed, 0.01, 0.5, 0.01, 0.01, 0.05)
    g.set_varcomp((0.6, 0.4, 0.4, 0.2, 0.2))
    return g, c, rng


def _all_variants_equal(g, x, prec):
    """V x with the tuned variant, then with every candidate forced; returns the tuned
    result after asserting that every candidate reproduces it bit for bit."""
    g.force_variant(-1)
    ref = g.computeVx_multi(x, precision=prec)
    counts = g.variant_counts()
    assert min(counts) >= 1
    for i in range(max(counts)):
        g.force_variant(i)
        y = g.computeVx_multi(x, precision=prec)
        assert np.array_equal(y, ref), f"candidate {i} of {counts}: max diff {np.abs(y - ref).max():.3e}"
    g.force_variant(-1)
    return ref, counts


def test_every_candidate_bit_identical_and_vs_valu(gpu_ctx_factory):
    n, m = 20000, 40000
    _say("variants", n, m)
    g, c, rng = _cohort_ctx(gpu_ctx_factory, n, m, 2.0, 7)
    for prec, k in CASES:
        _say("variants", "fp64" if prec == G.PREC_FP64 else "mixed", "k", k)
        x = rng.standard_normal((c.nrec, k))
        y, counts = _all_variants_equal(g, x, prec)
        y_valu = g.computeVx_multi(x, precision=G.PREC_FP64_VALU)
        e = rel(y, y_valu)
        assert e <= BOUND[prec], (prec, k, e)
        # the candidate lists reach the large variants at this cohort
        if prec == G.PREC_FP64 and k == 1:
            assert counts[0] >= 10 and counts[1] >= 10
        if prec == G.PREC_MIXED and k == 32:
            assert counts[0] >= 9 and counts[1] >= 9
    g.close()


def test_grm_builder_every_candidate(gpu_ctx_factory):
    """The builder's pass 2 runs on rows r0 .. n - 1 from multiples of 512 (4 digits: 128
    digit columns, and 32 for the last 8 columns of 9000); every forced candidate gives the
    same GRM bit for bit (a forced index wraps modulo the shape's candidate count, so the
    indices up to the largest count cover every candidate of every launch shape)."""
    n, m = 9000, 6000
    _say("builder variants", n, m)
    ref = None
    most = max(G.plan_replay(128, m, n, r0)["cand2"] for r0 in range(0, n, 512))
    for i in [-1] + list(range(most)):
        g, c, _ = _cohort_ctx(gpu_ctx_factory, n, m, 1.0, 11)
        g.force_variant(i)
        g.grm_from_geno()
        lt = g.grm_lt()
        g.close()
        if ref is None:
            ref = lt
        else:
            assert np.array_equal(lt, ref), f"candidate {i}"


def test_scale_shape_fp64_and_mixed_vs_valu(gpu_ctx_factory):
    """N = 200k, M = 500k, ~1.6M records (the scale workload): FP64 and mixed V x at k = 1
    and k = 32, every candidate forced, against the FP64 vector-ALU path."""
    n, m = 200000, 500000
    _say("scale variants", n, m)
    g, c, rng = _cohort_ctx(gpu_ctx_factory, n, m, 7.0, 2026)
    for prec, k in [(G.PREC_FP64, 1), (G.PREC_MIXED, 1), (G.PREC_MIXED, 32), (G.PREC_FP64, 32)]:
        _say("scale", "fp64" if prec == G.PREC_FP64 else "mixed", "k", k)
        x = rng.standard_normal((c.nrec, k))
        y, _ = _all_variants_equal(g, x, prec)
        e = rel(y, g.computeVx_multi(x, precision=G.PREC_FP64_VALU))
        _say("scale", prec, k, "rel err vs FP64 VALU", f"{e:.3e}")
        assert e <= BOUND[prec], (prec, k, e)
    g.close()
