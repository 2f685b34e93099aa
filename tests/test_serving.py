"""Serving loop (continuous batching) on the GPU: requests with their own individual
subsets, records and variance components join a running block PCG between
iterations. Each result must equal the same request served alone (the per-column
arithmetic does not see the other columns) and match the FP64 oracle of that
request's own system (its records, the subset's genotypes, the cohort's allele
frequencies)."""
import numpy as np
import pytest

import globe_amd as G
from globe_amd import serving, synth
from helpers import rel

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def cohort():
    return synth.make_geno_case(n=400, m=900, lam=3, seed=11, miss=0.01)


def _requests(n, count, seed):
    rng = np.random.default_rng(seed)
    return [serving.make_request(rng, n, 20, n, 1.0, 5.0) for _ in range(count)]


def _device(d, ctx_factory):
    g = ctx_factory()
    c = d.cohort
    g.set_cohort(c.indiv, c.times, c.X)
    g.prepGenotypeInput(d.bed_rows, 0.01, 0.2)
    g.set_varcomp(d.theta)
    return g


def _serve(g, reqs, *, slots, precision, tol, stagger=True, cap=None):
    # the slot capacity fixes the row partition of the reductions: alone and in the mix
    # use the same one, so the two runs do identical arithmetic
    cap = cap or max(r[0].shape[0] for r in reqs)
    g.serve_open(slots=slots, max_records=cap, tol=tol, max_iter=500, precision=precision)
    ids = {}
    pending = list(range(len(reqs)))
    first = pending[: len(reqs) // 2] if stagger else pending
    for q in first:
        ids[g.serve_submit(*reqs[q])] = q
    rest = pending[len(first):]
    out = [None] * len(reqs)
    info = [None] * len(reqs)
    steps = 0
    while ids or rest:
        steps += 1
        if rest and steps % 3 == 0:  # later arrivals join a running block
            q = rest.pop(0)
            ids[g.serve_submit(*reqs[q])] = q
        g.serve_step()
        for rid in g.serve_finished():
            x, st = g.serve_take(rid)
            q = ids.pop(rid)
            out[q], info[q] = x, st
        assert steps < 5000
    g.serve_close()
    return out, info


def _oracle_for(d, req):
    import oracle as O
    indiv, times, y, theta = req
    ids, sub = np.unique(indiv, return_inverse=True)
    codes = synth.unpack_codes(d.bed_rows[d.keep], d.cohort.n)[:, ids]
    dos = np.select([codes == 0, codes == 2, codes == 3], [2, 1, 0], default=-1).astype(np.int8)
    rows = synth.pack_bed(dos)
    X = np.column_stack([np.ones(indiv.shape[0]), times])
    return O.Oracle(sub.astype(np.int32), times, X, theta, geno_rows=rows, freq=d.freq, n=ids.shape[0])


def test_serving_mix_equals_alone_and_oracle(cohort, gpu_ctx_factory):
    _say("serve-fp64")
    d = cohort
    reqs = _requests(d.cohort.n, 10, seed=3)
    g = _device(d, gpu_ctx_factory)
    cap = max(r[0].shape[0] for r in reqs)
    mix, info = _serve(g, reqs, slots=4, precision=G.PREC_FP64, tol=1e-10, cap=cap)
    assert all(i["converged"] for i in info)
    for q in (0, 4, 9):
        alone, _ = _serve(g, [reqs[q]], slots=4, precision=G.PREC_FP64, tol=1e-10, stagger=False, cap=cap)
        assert rel(mix[q], alone[0]) <= 1e-14, q
        orc = _oracle_for(d, reqs[q])
        xo, _ = orc.pcg(np.asarray(reqs[q][2])[:, None], tol=1e-10, rule=2, stagnation=False)
        assert rel(mix[q], xo[:, 0]) < 1e-7, q


def test_serving_iteration_cap(cohort, gpu_ctx_factory):
    """Requests that reach max_iter retire unconverged after exactly max_iter iterations
    (their column is deactivated and leaves the block), holding the iterate the oracle's
    PCG reaches in as many iterations; in a block they equal the same request alone."""
    _say("serve-cap")
    d = cohort
    reqs = _requests(d.cohort.n, 6, seed=9)
    g = _device(d, gpu_ctx_factory)
    cap = max(r[0].shape[0] for r in reqs)

    def run(rs):
        g.serve_open(slots=3, max_records=cap, tol=1e-15, max_iter=4, precision=G.PREC_FP64)
        ids = {g.serve_submit(*r): q for q, r in enumerate(rs)}
        out, info = [None] * len(rs), [None] * len(rs)
        for _ in range(1000):
            g.serve_step()
            for rid in g.serve_finished():
                x, st = g.serve_take(rid)
                q = ids.pop(rid)
                out[q], info[q] = x, st
            if not ids:
                break
        g.serve_close()
        assert not ids
        return out, info

    mix, info = run(reqs)
    for q, st in enumerate(info):
        assert not st["converged"] and st["iters"] == 4, (q, st)
    for q in (0, 3, 5):
        alone, _ = run([reqs[q]])
        assert rel(mix[q], alone[0]) <= 1e-14, q
        orc = _oracle_for(d, reqs[q])
        xo, so = orc.pcg(np.asarray(reqs[q][2])[:, None], tol=1e-15, max_iter=4, rule=2, stagnation=False)
        assert so["iters"] == 4
        assert rel(mix[q], xo[:, 0]) < 1e-9, q


def test_serving_mixed_precision_within_1e6(cohort, gpu_ctx_factory):
    _say("serve-mixed")
    d = cohort
    reqs = _requests(d.cohort.n, 8, seed=5)
    g = _device(d, gpu_ctx_factory)
    x64, _ = _serve(g, reqs, slots=8, precision=G.PREC_FP64, tol=1e-10)
    xm, info = _serve(g, reqs, slots=3, precision=G.PREC_MIXED, tol=1e-10)
    assert all(i["converged"] for i in info)
    for q in range(len(reqs)):
        assert rel(xm[q], x64[q]) < 1e-6, q


def test_serving_rejects_bad_requests(cohort, gpu_ctx_factory):
    _say("serve-errors")
    d = cohort
    g = _device(d, gpu_ctx_factory)
    g.serve_open(slots=2, max_records=50, tol=1e-8)
    th = (0.5, 0.4, 0.3, 0.0, 0.2)
    with pytest.raises(G.GlobeError):  # more records than a slot holds
        g.serve_submit(np.zeros(100, np.int32), np.zeros(100), np.ones(100), th)
    with pytest.raises(G.GlobeError):  # individual outside the cohort
        g.serve_submit(np.array([0, 1, 10_000], np.int32), np.zeros(3), np.ones(3), th)
    with pytest.raises(G.GlobeError):  # records of individual 0 not contiguous
        g.serve_submit(np.array([0, 1, 0], np.int32), np.zeros(3), np.ones(3), th)
    with pytest.raises(G.GlobeError):  # unknown / unfinished request
        g._check(g.L.globe_serve_take(g.h, 12345, None, None))
    g.serve_close()


def test_serving_admits_by_hbm_budget(cohort, gpu_ctx_factory):
    """A budget whose record pool holds one request (80 bytes per pool element: 6 record
    vectors of 8 bytes + the record structure; a request takes nrec + 1 elements rounded
    up to 64) admits one at a time, FIFO, however many slots are free; a request larger
    than the budget is refused at submission; each result equals the same request served
    with an unlimited budget."""
    _say("serve-budget")
    d = cohort
    n = d.cohort.n
    rng = np.random.default_rng(9)
    # whole-cohort requests of similar size: no two fit a budget of the largest one
    reqs = [serving.make_request(rng, n, n, n, 2.0, 2.2) for _ in range(5)]
    sizes = sorted(r[0].shape[0] for r in reqs)
    assert sizes[0] + sizes[1] > sizes[-1]
    cap = sizes[-1]
    g = _device(d, gpu_ctx_factory)
    free, _ = _serve(g, reqs, slots=4, precision=G.PREC_FP64, tol=1e-10, stagger=False, cap=cap)
    per_rec = 80

    def span(nrec):
        return (nrec + 1 + 63) // 64 * 64

    assert span(sizes[0]) + span(sizes[1]) > span(cap)
    g.serve_open(slots=4, max_records=cap, hbm_budget=per_rec * span(cap), tol=1e-10, max_iter=500,
                 precision=G.PREC_FP64)
    ids = [g.serve_submit(*r) for r in reqs]
    order, out = [], {}
    for _ in range(5000):
        fin, active, queued = g.serve_step()
        assert active <= 1
        for rid in g.serve_finished():
            x, st = g.serve_take(rid)
            out[rid] = x
            order.append(rid)
        if len(out) == len(reqs):
            break
    assert order == ids  # first in, first admitted, first out
    for q, rid in enumerate(ids):
        assert rel(out[rid], free[q]) <= 1e-14, q
    g.serve_close()
    g.serve_open(slots=4, max_records=cap, hbm_budget=per_rec * 64, tol=1e-10)
    with pytest.raises(G.GlobeError):  # larger than the whole budget
        g.serve_submit(*reqs[0])
    g.serve_close()
    with pytest.raises(G.GlobeError):  # a budget that holds no request at all
        g.serve_open(slots=4, max_records=cap, hbm_budget=per_rec * 10, tol=1e-10)


def test_serving_pool_admits_by_request_size(cohort, gpu_ctx_factory):
    """Requests take pool ranges of their own size, not a slot of the largest one: with a
    pool that holds one large request, several small requests run together while the
    large one waits its turn (FIFO), and every result equals the same request served
    alone."""
    _say("serve-pool")
    d = cohort
    n = d.cohort.n
    rng = np.random.default_rng(21)
    big = serving.make_request(rng, n, n, n, 3.0, 3.0)
    smalls = [serving.make_request(rng, n, 30, 40, 1.0, 2.0) for _ in range(4)]
    reqs = smalls[:2] + [big] + smalls[2:]
    cap = big[0].shape[0]

    def span(nrec):
        return (nrec + 1 + 63) // 64 * 64

    assert 4 * max(span(r[0].shape[0]) for r in smalls) <= span(cap)
    g = _device(d, gpu_ctx_factory)
    alone = [_serve(g, [r], slots=4, precision=G.PREC_FP64, tol=1e-10, stagger=False, cap=cap)[0][0]
             for r in reqs]
    g.serve_open(slots=4, max_records=cap, hbm_budget=80 * span(cap), tol=1e-10, max_iter=500,
                 precision=G.PREC_FP64)
    ids = [g.serve_submit(*r) for r in reqs]
    out, most_active, order = {}, 0, []
    for _ in range(5000):
        _, active, _ = g.serve_step()
        most_active = max(most_active, active)
        for rid in g.serve_finished():
            out[rid] = g.serve_take(rid)[0]
            order.append(rid)
        if len(out) == len(reqs):
            break
    g.serve_close()
    assert most_active >= 2  # the two small requests ran side by side
    assert order.index(ids[2]) > order.index(ids[1])  # the large one waited for room
    for q, rid in enumerate(ids):
        assert rel(out[rid], alone[q]) <= 1e-14, q


def test_serving_submissions_from_an_intake_thread(cohort, gpu_ctx_factory):
    """globe_serve_submit runs on an intake thread (serving.replay: Poisson arrivals
    submitted from a second host thread while this one steps the loop): every request
    comes back once, converged, equal to the same request served alone to 1e-14, and a
    request the loop rejects fails on the intake thread with its own message."""
    _say("serve-intake-thread")
    d = cohort
    reqs = _requests(d.cohort.n, 24, seed=13)
    cap = max(r[0].shape[0] for r in reqs)
    g = _device(d, gpu_ctx_factory)
    g.serve_open(slots=6, max_records=cap, tol=1e-10, max_iter=500, precision=G.PREC_FP64)
    sols, st = serving.replay(g, reqs, rate=2000.0, seed=1)
    g.serve_close()
    assert all(x is not None for x in sols)
    assert st["iters"].min() >= 1 and st["steps"] >= 1
    for q in (0, 7, 23):
        alone, info = _serve(g, [reqs[q]], slots=6, precision=G.PREC_FP64, tol=1e-10, stagger=False, cap=cap)
        assert info[0]["converged"]
        assert rel(sols[q], alone[0]) <= 1e-14, q
    # an invalid request raised on the intake thread reaches the stepping thread
    g.serve_open(slots=2, max_records=cap, tol=1e-8)
    bad = (np.array([0, 1, 0], np.int32), np.zeros(3), np.ones(3), (0.5, 0.4, 0.3, 0.0, 0.2))
    with pytest.raises(G.GlobeError, match="contiguous"):
        serving.replay(g, [reqs[0], bad], rate=2000.0, seed=2)
    g.serve_close()


def test_serving_row_gather_equals_all_rows(gpu_ctx_factory):
    """Requests over small subsets of a 3,000-individual cohort: the individual-major
    pass formed only for the requests' individuals (gathered rows, padded to 512) gives
    every request bit for bit the solution of the pass over every row, with the tuned
    thin shape and with every thin candidate forced; profiling shows the gathered pass
    reading fewer rows."""
    _say("serving row gather")
    d = synth.make_geno_case(n=3000, m=2000, lam=3, seed=21, miss=0.01)
    rng = np.random.default_rng(5)
    reqs = [serving.make_request(rng, 3000, 40, 400, 1.0, 4.0) for _ in range(8)]
    out = {}
    for on in (False, True):
        g = _device(d, gpu_ctx_factory)
        g.set_row_gather(on)
        g.profile(True)
        out[on], info = _serve(g, reqs, slots=4, precision=G.PREC_MIXED, tol=1e-8)
        ks = g.kernel_stats(1)
        g.profile(False)
        full_bytes = 3000 * ((g.m + 3) // 4)
        if on:
            assert 0 < ks["bytes"] < full_bytes, (ks["bytes"], full_bytes)
            counts = g.variant_counts()
        g.close()
        assert all(st["converged"] for st in info)
    for q in range(len(reqs)):
        assert np.array_equal(out[True][q], out[False][q]), q
    for v in range(max(counts[1], 1)):
        _say("serving row gather, forced variant", v)
        g = _device(d, gpu_ctx_factory)
        g.force_variant(v)
        x, _ = _serve(g, reqs, slots=4, precision=G.PREC_MIXED, tol=1e-8)
        g.close()
        for q in range(len(reqs)):
            assert np.array_equal(x[q], out[False][q]), (v, q)


def test_serving_row_gather_ragged_pitch(gpu_ctx_factory):
    """A cohort whose individual-major pitch (640 bytes) is not a multiple of the gathered
    pass's 256 bytes per workgroup, so its last workgroup reads clamped pieces past the
    pitch and drops their SNPs: gathered and full passes agree bit for bit in mixed
    precision (the last packed byte ragged too) and in FP64 with two requests in the
    block, and an FP64 request (7 digits, gathered alone in the block) matches the
    oracle of its own system."""
    _say("serving row gather, ragged pitch")
    d = synth.make_geno_case(n=2500, m=2302, lam=3, seed=23, miss=0.01)
    rng = np.random.default_rng(8)
    reqs = [serving.make_request(rng, 2500, 30, 300, 1.0, 4.0) for _ in range(6)]
    out = {}
    for on in (False, True):
        g = _device(d, gpu_ctx_factory)
        pitch = ((g.m + 3) // 4 + 127) // 128 * 128  # individual-major row bytes
        assert pitch % 256 != 0 and g.m % 4 != 0, (g.m, pitch)
        g.set_row_gather(on)
        g.profile(True)
        out[on], info = _serve(g, reqs, slots=3, precision=G.PREC_MIXED, tol=1e-8)
        ks = g.kernel_stats(1)
        g.profile(False)
        if on:
            assert 0 < ks["bytes"] < 2500 * ((g.m + 3) // 4), ks
        g.close()
        assert all(st["converged"] for st in info)
    for q in range(len(reqs)):
        assert np.array_equal(out[True][q], out[False][q]), q
    # FP64 with two requests in the block: 14 digit columns, still one gathered group
    x64 = {}
    for on in (False, True):
        g = _device(d, gpu_ctx_factory)
        g.set_row_gather(on)
        x64[on], info = _serve(g, reqs[:2], slots=2, precision=G.PREC_FP64, tol=1e-10, stagger=False)
        g.close()
        assert all(st["converged"] for st in info)
    for q in range(2):
        assert np.array_equal(x64[True][q], x64[False][q]), q
    g = _device(d, gpu_ctx_factory)
    g.profile(True)
    x, info = _serve(g, [reqs[2]], slots=2, precision=G.PREC_FP64, tol=1e-10, stagger=False)
    ks = g.kernel_stats(1)
    g.close()
    assert 0 < ks["bytes"] < 2500 * ((g.m + 3) // 4), ks
    assert info[0]["converged"]
    orc = _oracle_for(d, reqs[2])
    xo, _ = orc.pcg(np.asarray(reqs[2][2])[:, None], tol=1e-10, rule=2, stagnation=False)
    assert rel(x[0], xo[:, 0]) < 1e-7
