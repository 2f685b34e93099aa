"""Host replay of the packed-genotype contraction's launch plans (CPU, no GPU).

The tuner times every candidate variant on live data, so every candidate of a shape
runs on the GPU at least once, not only the winner. plan_replay walks every block of
every candidate of both passes through the kernels' own index functions
(globe_amd/csrc/contract_plan.h) and fails on the first read outside the padded packed
layout, digit access outside the digit matrix, partial-sum store or epilogue read
outside the split-partial buffer that the library allocates, or partial slot left
unwritten. The shapes are those the benchmarks and the GRM builder run:
N = 200k / M = 500k (the scale workload), N = 40k / M = 80k with the builder's
512-multiple row offsets (the AI-REML cohort), N = 200k from 20k SNPs (the --grm
path's builder), plus small and ragged shapes of the tests.
"""
import pytest

import globe_amd

# digit-column counts: the geometry depends on J only through the fragment count (thin:
# J <= 16; wide: 1 to 4 fragments of 32), so one count per class covers them all
J_ALL = (7, 28, 63, 91, 128)
# resident thin blocks per CU: the kernels' LDS admits 1 or 2; the replay takes both
BPC = ((1,) * 8, (2,) * 8, (2, 1, 1, 1, 1, 1, 2, 1))


def _lds_bpc():
    """Blocks per CU each thin shape's LDS ring admits (160 KB per CU)."""
    shapes = [(4, 4, 3, 1), (8, 2, 3, 2), (8, 2, 4, 2), (16, 1, 3, 2), (16, 1, 4, 2), (8, 1, 5, 2),
              (4, 4, 4, 2), (8, 1, 6, 2)]
    out = []
    for w, rt, ns, ss in shapes:
        slot = w * rt * ss * 1024 + 4 * ss * 1024
        out.append(max(1, min(2, (160 * 1024) // (ns * slot))))
    return tuple(out)


def _replay_all(m, n, r0s=(0,), js=J_ALL, ncus=(256,)):
    seen_s = set()
    for j in js:
        for bpc in BPC + (_lds_bpc(),):
            for ncu in ncus:
                for r0 in r0s:
                    st = globe_amd.plan_replay(j, m, n, r0=r0, ncu=ncu, thin_bpc=bpc)
                    assert st["row_margin1"] <= 0 and st["row_margin2"] <= 0
                    seen_s.add((st["wide"], st["cand1"], st["cand2"]))
    return seen_s


def test_replay_scale_workload_all_candidates():
    """N = 200k, M = 500k (499,939 kept): the headline's thin and wide candidates."""
    _replay_all(499939, 200000)


def _builder_offsets(n, stride):
    """Row offsets r0 of the builder's launches: multiples of 512 (every stride-th one,
    and all of the last 16, where the fewest rows admit the most splits)."""
    last = n // 512 * 512 if n % 512 else n - 512
    return tuple(sorted(set(range(0, n, stride)) | set(range(max(0, last - 15 * 512), last + 1, 512))))


def test_replay_aireml_cohort_builder_offsets():
    """N = 40k, M = 80k: the GRM builder contracts rows r0 .. n - 1 from multiples of 512
    below each tile column (fewer rows admit more splits: the shape of the one fault of the
    first builder), 32 columns of 4 digits (and 32 digit columns for an 8-column rest)."""
    _replay_all(79990, 40000, r0s=_builder_offsets(40000, 4096), js=(128, 32))


def test_replay_grm_path_builder():
    """N = 200k from 20k SNPs (the --grm path at scale): builder offsets up to 199,168."""
    _replay_all(20000, 200000, r0s=_builder_offsets(200000, 4096 * 6), js=(128, 32))


@pytest.mark.parametrize("m,n", [(1, 1), (3, 5), (900, 301), (2000, 400), (513, 1025), (40000, 20000),
                                 (4097, 70001)])
def test_replay_small_and_ragged(m, n):
    _replay_all(m, n, r0s=(0,) if n < 4096 else (0, 4096), ncus=(256, 80))


def test_replay_candidate_sets_reach_the_large_variants():
    """The N = 20k / M = 40k cohort of the forced-variant GPU test has S > 8 split-K
    candidates and multi-piece stream-K candidates, so that test runs those paths."""
    st = globe_amd.plan_replay(128, 40000, 20000)
    assert st["wide"] and st["cand1"] >= 9 and st["cand2"] >= 9
    st = globe_amd.plan_replay(7, 40000, 20000)
    assert not st["wide"] and st["cand1"] >= 10


def test_replay_detects_an_undersized_buffer():
    """The replay is not vacuous: the digit matrix and the split-partial buffer are
    exactly as large as the farthest access of some candidate, and the packed layouts
    exceed the farthest row read by the replay's reported margin, so taking 16 bytes
    more than that margin off the layouts, 16 off the digit matrix or 4 off the
    partial buffer makes the replay fail."""
    for j, m, n in ((128, 40000, 20000), (7, 40000, 20000), (125, 79990, 40000)):
        st = globe_amd.plan_replay(j, m, n)
        slack = -max(st["row_margin1"], st["row_margin2"])
        for shrink, what in (((slack + 16, 0, 0), "packed-row read"), ((0, 16, 0), "digit"),
                             ((0, 0, 4), "partial")):
            with pytest.raises(AssertionError, match=what):
                globe_amd.plan_replay(j, m, n, shrink=shrink)
        globe_amd.plan_replay(j, m, n, shrink=(slack, 0, 0))


def test_replay_rejects_bad_shapes():
    with pytest.raises(AssertionError):
        globe_amd.plan_replay(129, 40000, 20000)
    with pytest.raises(AssertionError):
        globe_amd.plan_replay(128, 40000, 20000, r0=100)


def _gather_sizes(n):
    """Listed-individual counts of the gathered pass: 512 up to the cohort padded to 512
    (every 512-multiple for small cohorts, a spread plus the largest for large ones)."""
    top = (n + 511) // 512 * 512
    sizes = range(512, top + 1, 512)
    if len(sizes) > 40:
        sizes = sorted(set(list(sizes[:8]) + list(sizes[::max(1, len(sizes) // 24)]) + list(sizes[-4:])))
    return sizes


@pytest.mark.parametrize("m,n", [(499939, 200000), (2000, 3000), (2101, 2500), (2302, 2500), (1, 1), (5, 513), (900, 400),
                                 (4097, 70001), (79990, 40000)])
def test_replay_gathered_pass1(m, n):
    """k_geno_tg over every listed-individual count the serving loop can pass: the
    workgroups tile the individual-major pitch (a last one past the pitch reads clamped
    pieces), the splits tile the stages, each stage row's LDS slots are filled once with
    the pieces the transposing reads expect, and every SNP < m of every split is stored
    once inside the partial buffer.
    The gathered pass 2 (a thin candidate over the ng listed rows) is every candidate of
    a thin product on ng rows, replayed for the same counts (at the scale cohort for one
    CU count and residency only: each replay there also walks pass 1's 500k rows)."""
    seen = set()
    for ng in _gather_sizes(n):
        for ncu in (256, 80):
            for bpc in BPC:
                st = globe_amd.plan_replay_tg(m, n, ng, ncu=ncu, thin_bpc=bpc)
                assert st["row_margin"] <= 0
                seen.add((st["workgroups"], st["splits"]))
                if m < 100000 or (ncu, bpc) == (256, BPC[0]):
                    st2 = globe_amd.plan_replay(7, m, ng, ncu=ncu, thin_bpc=bpc)
                    assert not st2["wide"] and st2["row_margin2"] <= 0
    if n >= 40000:
        assert max(p for _, p in seen) > 1, seen  # split launches are replayed too


def test_replay_gathered_pass1_detects_an_undersized_buffer():
    """Taking 16 bytes more than the reported margin off the packed layout, or 16 bytes
    off a digit matrix used to its last byte, fails the replay; a smaller partial
    buffer gives fewer splits, down to one slot of m x 16 sums, and one short of that
    fails."""
    m, n, ng = 2101, 2500, 2560
    st = globe_amd.plan_replay_tg(m, n, ng)
    with pytest.raises(AssertionError, match="packed-row read"):
        globe_amd.plan_replay_tg(m, n, ng, shrink=(-st["row_margin"] + 16, 0, 0))
    globe_amd.plan_replay_tg(m, n, ng, shrink=(-st["row_margin"], 0, 0))
    # 2,048 individuals and a 256-byte pitch: the digit matrix is 16 x 2,048 bytes
    st = globe_amd.plan_replay_tg(900, 2048, 2048)
    assert st["digit_bytes"] == 16 * 2048
    with pytest.raises(AssertionError, match="digit"):
        globe_amd.plan_replay_tg(900, 2048, 2048, shrink=(0, 16, 0))
    m, n, ng = 499939, 200000, 65536
    st = globe_amd.plan_replay_tg(m, n, ng)
    assert st["splits"] > 1
    slot = m * 16 * 4
    fewer = globe_amd.plan_replay_tg(m, n, ng, shrink=(0, 0, st["part_bytes"] - slot))
    assert fewer["splits"] == 1
    with pytest.raises(AssertionError, match="partial"):
        globe_amd.plan_replay_tg(m, n, ng, shrink=(0, 0, st["part_bytes"] - slot + 4))


def test_replay_gathered_pass1_rejects_bad_counts():
    for ng in (0, 100, 4096):
        with pytest.raises(AssertionError):
            globe_amd.plan_replay_tg(2000, 3000, ng)
