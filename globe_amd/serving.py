"""Serving-loop driver: synthetic solve requests against a resident cohort, replayed
with Poisson arrivals in real time through globe_serve_* (continuous batching).

A request is a trait on an individual subset of the resident cohort: its own
variance components, its own longitudinal records (1 + Poisson(lambda) per
individual, times U(0, 10)) and its own right-hand side. The driver submits each
request at its arrival time, steps the loop (one admission + block-PCG iteration
per step) and collects results; latency = completion - arrival on the host clock.
"""
from __future__ import annotations

import time

import numpy as np


def draw_theta(rng):
    """sig2g U(0.1, 0.6), sig2e U(0.3, 0.8), sig2b0 U(0.1, 0.5), sig2b1 U(0.05, 0.3),
    correlation of the random intercept and slope U(-0.5, 0.5)."""
    sg = rng.uniform(0.1, 0.6)
    se = rng.uniform(0.3, 0.8)
    sb0 = rng.uniform(0.1, 0.5)
    sb1 = rng.uniform(0.05, 0.3)
    rho = rng.uniform(-0.5, 0.5)
    return (se, sg, sb0, rho * np.sqrt(sb0 * sb1), sb1)


def make_request(rng, n_cohort, size_lo, size_hi, lam_lo, lam_hi, u_pool=None):
    """One request: (indiv, times, y, theta). Subset size log-uniform in [size_lo,
    size_hi]; records 1 + Poisson(lambda), lambda ~ U(lam_lo, lam_hi); y = u_i + b0_i +
    b1_i t + e with the request's components (u from a pool of genetic values scaled
    to the request's sig2g, or absent)."""
    size = int(np.exp(rng.uniform(np.log(size_lo), np.log(size_hi))))
    size = max(1, min(size, n_cohort))
    ids = np.sort(rng.choice(n_cohort, size=size, replace=False)).astype(np.int32)
    lam = rng.uniform(lam_lo, lam_hi)
    counts = 1 + rng.poisson(lam, size=size)
    indiv = np.repeat(ids, counts)
    nrec = indiv.shape[0]
    times = rng.uniform(0.0, 10.0, size=nrec).astype(np.float32).astype(np.float64)
    theta = draw_theta(rng)
    se, sg, sb0, sb01, sb1 = theta
    L = np.linalg.cholesky(np.array([[sb0, sb01], [sb01, sb1]]))
    b = rng.standard_normal((size, 2)) @ L.T
    bi = np.repeat(b, counts, axis=0)
    y = bi[:, 0] + bi[:, 1] * times + np.sqrt(se) * rng.standard_normal(nrec)
    if u_pool is not None:
        u = u_pool[rng.integers(len(u_pool))]
        y += np.sqrt(sg) * u[indiv]
    return indiv, times, y, theta


def replay(g, requests, rate, seed=0, intake_threads=4):
    """Submit requests at Poisson arrival times (rate per second) from intake threads
    (request q on thread q mod intake_threads; globe_serve_submit prepares each request on
    the host while the loop steps the GPU) and run the loop until all finish. Returns
    (solutions, stats) with per-request latency (completion - arrival), its parts (intake:
    arrival to queued, the request prepared on the host; queue: queued to admitted;
    service: admitted to done, on the loop's clock) and loop counters. A request the loop
    rejects stops the intake and raises here."""
    import threading
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(requests)))
    arrivals -= arrivals[0]
    sols = [None] * len(requests)
    lat = np.zeros(len(requests))
    iters = np.zeros(len(requests), dtype=np.int64)
    intake_s = np.zeros(len(requests))
    queue_s = np.zeros(len(requests))
    service_s = np.zeros(len(requests))
    idx = {}
    lock = threading.Lock()
    stop = threading.Event()
    failure = []
    nthreads = max(1, min(int(intake_threads), len(requests)))
    t0 = time.perf_counter()

    def intake(t):
        try:
            for q in range(t, len(requests), nthreads):
                wait = arrivals[q] - (time.perf_counter() - t0)
                if (wait > 0 and stop.wait(wait)) or stop.is_set():
                    return
                rid = g.serve_submit(*requests[q])
                intake_s[q] = (time.perf_counter() - t0) - arrivals[q]
                with lock:
                    idx[rid] = q
        except Exception as e:  # surfaced by the stepping thread
            failure.append(e)
            stop.set()

    threads = [threading.Thread(target=intake, args=(t,), daemon=True) for t in range(nthreads)]
    for th in threads:
        th.start()
    steps = 0
    occupancy = 0
    done = 0
    try:
        while done < len(requests):
            if failure:
                raise failure[0]
            _, active, queued = g.serve_step()
            if active:
                steps += 1
                occupancy += active
            for rid in g.serve_finished():
                with lock:
                    q = idx.pop(rid, None)
                if q is None:  # its id is not recorded yet: take it at the next step
                    continue
                x, st = g.serve_take(rid)
                sols[q] = x
                lat[q] = (time.perf_counter() - t0) - arrivals[q]
                iters[q] = st["iters"]
                queue_s[q] = st["t_admit"] - st["t_submit"]
                service_s[q] = st["t_done"] - st["t_admit"]
                done += 1
            if not active and not queued:
                time.sleep(2e-4)  # idle: leave the interpreter to the intake threads
    finally:
        # no submit may overlap the caller's next call (globe_serve_close)
        stop.set()
        for th in threads:
            th.join()
    wall = time.perf_counter() - t0
    return sols, {"latency_s": lat, "iters": iters, "wall_s": wall, "steps": steps,
                  "mean_active": occupancy / max(steps, 1), "intake_s": intake_s, "queue_s": queue_s,
                  "service_s": service_s}


def main(argv=None):
    """Replay synthetic requests on a synthetic resident cohort (genotypes generated in
    HBM): python -m globe_amd.serving --n 200000 --m 500000 --requests 200 --rate 1000.
    A 16-byte device copy before and after the replay marks its span in a kernel trace."""
    import argparse
    import json
    import globe_amd as G
    from globe_amd import synth
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--m", type=int, default=500_000)
    ap.add_argument("--lam", type=float, default=7.0)
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--rate", type=float, default=1000.0)
    ap.add_argument("--slots", type=int, default=32)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=20261206)
    a = ap.parse_args(argv)
    rng = np.random.default_rng(a.seed)
    g = G.Globe(0)
    c = synth.simulate_cohort(a.n, a.lam, rng)
    g.set_cohort(c.indiv, c.times, c.X)
    kept = g.synth_bed(a.m, seed=a.seed, f_lo=0.01, f_hi=0.5, miss=0.01, min_maf=0.01, max_miss=0.05)
    g.set_varcomp((0.6, 0.4, 0.4, 0.2, 0.2))
    u_pool = [np.sqrt(1.0 / kept) * g.geno_Wa(rng.standard_normal(kept)) for _ in range(4)]
    reqs = [make_request(rng, a.n, min(5_000, a.n), a.n, 2.0, 10.0, u_pool) for _ in range(a.requests)]
    cap = max(r[0].shape[0] for r in reqs)
    g.serve_open(slots=a.slots, max_records=cap, tol=a.tol, max_iter=1000, precision=G.PREC_MIXED)
    marker = g.alloc(256)
    g._check(g.L.globe_dev_copy(g.h, marker.ptr, marker.ptr + 128, 16))
    _, st = replay(g, reqs, a.rate, seed=a.seed)
    g._check(g.L.globe_dev_copy(g.h, marker.ptr, marker.ptr + 128, 16))
    g.serve_close()
    lat = st["latency_s"]
    print(json.dumps({"requests": a.requests, "rate": a.rate, "wall_s": st["wall_s"],
                      "solves_per_s": a.requests / st["wall_s"], "p50_s": float(np.percentile(lat, 50)),
                      "p99_s": float(np.percentile(lat, 99)), "steps": st["steps"],
                      "mean_active": st["mean_active"]}), flush=True)
    g.close()


if __name__ == "__main__":
    main()
