"""Benchmark: PCG solves/s of the packed-genotype LGMM solve on one MI355X.

Workload (the metric's configuration): N = 200,000 individuals, M = 500,000 SNPs
in PLINK 2-bit form (p ~ U(0.01, 0.5), 1% missing, QC minMAF 0.01 /
maxMissingRate 0.05), 1 + Poisson(7) records per individual (~1.6M), covariates
intercept / sex / age + TIME, variance components (0.6, 0.4, 0.4, 0.2, 0.2).
The genotypes are generated directly in HBM (25 GB packed, plus the
individual-major copy); y follows the LGMM recipe with u = sqrt(sig2g / M) W a.

A step is one block PCG solve of 32 right-hand sides (y + 31 Rademacher probes)
in mixed precision (int8 matrix cores, 4 digits), per-column convergence at
tol 1e-6. value = converged columns per second over the timed steps. Beside it:
the FP64 single-RHS solve of y (7-digit contraction) -- solves/s and the error of
the mixed solution of y against it -- the roofline of the dominant kernel from HIP
events on the library stream, and the CPU baseline (the reference's FP32 OpenMP
operator restated in oracle/, timed on a bounded SNP sample and extrapolated).
"""
from __future__ import annotations

import argparse
import json
import os
import select
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0
PEAK_INT8_TOPS = 5000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


MARK = False


def mark(g):
    """With --mark: a 16-byte device copy (k_copy16) that delimits a leg in a kernel trace."""
    if MARK:
        buf = g.alloc(256)
        g._check(g.L.globe_dev_copy(g.h, buf.ptr, buf.ptr + 128, 16))
        buf.free()


def build_cohort(g, n, m, lam, seed):
    from globe_amd import synth
    rng = np.random.default_rng(seed)
    c = synth.simulate_cohort(n, lam, rng)
    g.set_cohort(c.indiv, c.times, c.X)
    t0 = time.time()
    kept = g.synth_bed(m, seed=seed, f_lo=0.01, f_hi=0.5, miss=0.01, min_maf=0.01, max_miss=0.05)
    log(f"staged {m} SNPs x {n} individuals in HBM ({kept} kept) in {time.time() - t0:.1f}s")
    theta = (0.6, 0.4, 0.4, 0.2, 0.2)
    g.set_varcomp(theta)
    a = rng.standard_normal(kept)
    u = np.sqrt(theta[1] / kept) * g.geno_Wa(a)
    y = synth.simulate_phenotype(c, u, theta, rng)
    return c, y, kept


def measured_traffic(role):
    """HBM bytes per launch of the dominant kernel from the committed rocprof summary
    (calibrated FETCH_SIZE + WRITE_SIZE; profiles/k_geno_mfma_n200k_m500k.md has the
    commands): role "wide" (32-RHS contraction) or "thin" (1-RHS pass). None when no
    summary of that kernel is present."""
    path = os.path.join(ROOT, "profiles", "k_geno_mfma_n200k_m500k.traffic.json")
    try:
        with open(path) as f:
            tr = json.load(f)
        v = tr[role]
        return v["fetch_bytes"] + v["write_bytes"], f"{v['kernel']} / {v['write_kernel']}"
    except (OSError, ValueError, KeyError):
        return None, None


def cpu_baseline(g, c, kept, iters, sample_snps, threads):
    """FP32 OpenMP restatement of computeVx_multi (oracle/) on the first sample_snps kept
    SNPs at full N, k = 1; per-V.x time scaled by M / sample and by the FP64 iteration
    count to a solves/s figure."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    rows = g.geno_rows(0, sample_snps)
    freq = g.allele_freqs()[:sample_snps]
    orc = O.Oracle(c.indiv, c.times, c.X, (0.6, 0.4, 0.4, 0.2, 0.2), geno_rows=rows, freq=freq, n=c.n)
    x = np.random.default_rng(0).standard_normal(c.nrec).astype(np.float32)
    orc.vx_ref32(x)  # warm
    reps = 2
    t0 = time.time()
    for _ in range(reps):
        orc.vx_ref32(x)
    t_sample = (time.time() - t0) / reps
    t_full = t_sample * kept / sample_snps
    return {"value": 1.0 / (t_full * iters), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{reps} V.x of the FP32 OpenMP restatement of computeVx_multi at N={c.n}, "
                      f"Nrec={c.nrec}, k=1 on {sample_snps} of {kept} SNPs ({t_sample:.2f} s each), "
                      f"scaled linearly to M and by the GPU FP64 iteration count ({iters}); extrapolated",
            "sec_per_vx_extrapolated": t_full}


def _reference_iteration_time(exe, c, y, rows, theta, threads, max_iters):
    """Seconds per PCG iteration of the reference runPCGsolver on cohort c with the given
    .bed rows: its --debug mode prints one line per iteration; stdout is a pseudo-terminal
    so each line arrives as it is printed. The run is stopped after max_iters lines (or
    at convergence). Returns (seconds per iteration, iterations timed, threads, seconds
    of input parsing and set-up)."""
    import pty
    from globe_amd import synth
    tmp = tempfile.mkdtemp(prefix="globe_cpu_ref_")
    try:
        prefix = os.path.join(tmp, "sample")
        synth.write_plink(prefix, rows, c.n)
        synth.write_text_input(prefix + ".txt", c, y)
        cmd = [exe, "--input", prefix + ".txt", "--output", prefix + "_out", "--geno", prefix,
               "--minMAF", "0", "--maxMissingRate", "1", "--debug"]
        for f, v in zip(("--sig2e", "--sig2g", "--sig2b0", "--sigb01", "--sig2b1"), theta):
            cmd += [f, repr(float(v))]
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        master, slave = pty.openpty()
        t0 = time.perf_counter()
        proc = subprocess.Popen(cmd, stdout=slave, stderr=slave, env=env, close_fds=True)
        os.close(slave)
        stamps, nthr, t_start, buf, done = [], threads, None, b"", False
        deadline = time.perf_counter() + 180.0  # the reference gets 3 minutes per sample
        try:
            while len(stamps) < max_iters and not done:
                left = deadline - time.perf_counter()
                if left <= 0 or not select.select([master], [], [], left)[0]:
                    break
                try:
                    chunk = os.read(master, 65536)
                except OSError:
                    break
                if not chunk:
                    break
                now = time.perf_counter()
                buf += chunk
                *lines, buf = buf.split(b"\n")
                for ln in lines:
                    t = ln.decode(errors="replace")
                    if "OpenMP will utilize up to" in t:
                        nthr = int(t.split("up to")[1].split()[0])
                    elif "Starting PCG solver" in t:
                        t_start = now
                    elif "[runPCG_multi] Iter " in t:
                        stamps.append(now)
                    elif "Converged at iteration" in t:
                        stamps.append(now)
                        done = True
        finally:
            proc.kill()
            proc.wait(timeout=60)
            os.close(master)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    if t_start is None or len(stamps) < 3:
        return None
    # iteration 1 ends at stamps[0]; time the ones after it
    t_it = (stamps[-1] - stamps[0]) / (len(stamps) - 1)
    return t_it, len(stamps) - 1, nthr, t_start - t0


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_reference(g, c, y, kept, iters, sample_snps, threads, theta, max_iters=7):
    """The reference's own runPCGsolver (oracle/_ref, compiled from its sources by
    oracle/Makefile) timed on the host cores: the same cohort and phenotype with the
    first s kept SNPs written as a .bed, for several sample sizes s up to one of its
    50,000-SNP chunks (PCGsolver.cpp:1135; 2.5 GB at N = 200k, beyond the host caches).
    Its PCG solves [X | y] (p + 2 columns, FP32); the time per iteration is fitted by least
    squares as a + b s (record-space work plus genotype passes), the fit's residuals at the
    samples are reported, and it is evaluated at s = M and multiplied by the GPU FP64
    iteration count. Returns None when the binary is absent."""
    exe = os.path.join(ROOT, "oracle", "_ref", "runPCGsolver")
    if not os.access(exe, os.X_OK):
        return None
    pts = []
    for s in sample_snps:
        rows = g.geno_rows(0, s)[:, : (c.n + 3) // 4]
        r = _reference_iteration_time(exe, c, y, rows, theta, threads, max_iters)
        del rows
        if r is None:
            log(f"CPU baseline: the reference gave no timing at {s} SNPs")
            continue
        pts.append((s, *r))
        log(f"CPU baseline: reference at {s} SNPs: {r[0]:.3f} s per PCG iteration ({r[1]} timed, "
            f"{r[2]} threads, set-up {r[3]:.1f} s)")
    if len(pts) < 2:
        return None
    S = np.array([p[0] for p in pts], dtype=np.float64)
    T = np.array([p[1] for p in pts])
    A = np.column_stack([np.ones_like(S), S])
    (a, b), *_ = np.linalg.lstsq(A, T, rcond=None)
    resid = T - (a + b * S)
    t_full = a + b * kept
    nthr = pts[0][3]
    k = c.X.shape[1] + 1
    desc = "; ".join(f"{p[2]} iterations at {p[0]} SNPs, {p[1]:.3f} s each" for p in pts)
    return {"value": k / (t_full * iters), "unit": "solves/s", "cores": nthr, "kind": "reference",
            "cpu_model": _cpu_model(),
            "sample": f"the reference runPCGsolver (FP32, OpenMP, {nthr} threads, {_cpu_model()}) on N={c.n}, "
                      f"Nrec={c.nrec}, its {k}-column [X | y] solve: {desc} (input parsing and set-up not "
                      f"counted); per-iteration time a + b M fitted by least squares (a = {a:.3f} s, b = "
                      f"{b * 1e6:.3f} us per SNP; residuals {', '.join(f'{x:+.3f}' for x in resid)} s) and "
                      f"evaluated at M = {kept} ({t_full:.2f} s), times the GPU FP64 iteration count "
                      f"({iters}); {k} right-hand sides per solve; extrapolated",
            "fit": {"a_s": float(a), "b_s_per_snp": float(b), "snps": [int(x) for x in S],
                    "sec_per_iter": [float(x) for x in T], "residual_s": [float(x) for x in resid]},
            "sec_per_iter_extrapolated": t_full}


def serving_leg(g, c, kept, args):
    """Requests on the resident cohort replayed with Poisson arrivals through the serving
    loop (continuous batching into a 32-column mixed-precision block PCG, tol 1e-6):
    subsets log-U[5k, 200k], 1 + Poisson(lambda) records with lambda ~ U(2, 10), each
    with its own variance components; traits carry a genetic value from a pool of four
    W a draws."""
    from globe_amd import serving
    rng = np.random.default_rng(args.seed + 7)
    u_pool = [np.sqrt(1.0 / kept) * g.geno_Wa(rng.standard_normal(kept)) for _ in range(4)]
    t0 = time.time()
    reqs = [serving.make_request(rng, c.n, 5_000, min(200_000, c.n), 2.0, 10.0, u_pool)
            for _ in range(args.serve_requests)]
    cap = max(r[0].shape[0] for r in reqs)
    log(f"serving: {len(reqs)} requests generated in {time.time() - t0:.1f}s "
        f"(records {min(r[0].shape[0] for r in reqs)}..{cap})")
    import globe_amd as G
    # the record pool (request vectors and structure) gets a share of the HBM left free
    # by the resident cohort and the solver's workspaces
    free_hbm, total_hbm = g.device_memory()
    budget = int(args.serve_hbm_frac * free_hbm)

    def run(batch, rate):
        g.serve_open(slots=args.serve_slots, max_records=cap, hbm_budget=budget, tol=args.tol, max_iter=1000,
                     precision=G.PREC_MIXED)
        mark(g)
        _, st = serving.replay(g, batch, rate, seed=args.seed)
        mark(g)
        use = g.serve_usage()
        g.serve_close()
        lat = st["latency_s"]
        r = {"requests": len(batch), "arrival_rate_per_s": rate, "solves_per_s": len(batch) / st["wall_s"],
             "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
             "mean_iters": float(np.mean(st["iters"])), "mean_active_slots": st["mean_active"],
             "block_iterations": st["steps"], "wall_s": st["wall_s"],
             "pool_peak_bytes": use["peak"], "steps_waiting_for_hbm": use["waits"],
             # latency parts, p50 / p99: arrival to queued (host preparation on the intake
             # thread), queued to admitted, admitted to done
             "parts_s": {k: [float(np.percentile(st[k + "_s"], 50)), float(np.percentile(st[k + "_s"], 99))]
                         for k in ("intake", "queue", "service")}}
        log(f"serving at {rate:g}/s: {r['solves_per_s']:.1f} solves/s, p50 {r['p50_s']:.3f}s, "
            f"p99 {r['p99_s']:.3f}s, mean active {r['mean_active_slots']:.1f}, pool peak "
            f"{use['peak'] / 1e9:.2f} of {use['pool'] / 1e9:.2f} GB, waits for HBM {use['waits']}; "
            + ", ".join(f"{k} {v[0]:.3f}/{v[1]:.3f}s" for k, v in r["parts_s"].items()))
        return r

    out = run(reqs, args.serve_rate)
    # a tenth of the same requests offered 50x faster (a burst the loop cannot clear as it
    # arrives): throughput and latency with all slots busy
    out["saturated"] = run(reqs[: max(1, len(reqs) // 10)], 50.0 * args.serve_rate)
    out.update({"slots": args.serve_slots, "precision": "mixed", "tol": args.tol,
                "hbm_budget_bytes": budget, "hbm_free_bytes": free_hbm,
                "data": "synthetic requests (subset, records, components, trait) on the resident cohort"})
    return out


def aireml_cpu_baseline(g, c, theta, grm_launches, grm_cols, iters, sample_indiv, threads):
    """runAIREML's operator on the CPU: the FP32 OpenMP record-pair restatement of
    computeVx_multi_VC (varCompEst.cpp:322-352, oracle/) timed with 3 applications at
    k = 1 and 3 at k = 16 columns on the records of the first sample_indiv individuals
    (the leading block of the device's GRM), fitted as a + b k per application and scaled
    by (Nrec / Nrec_sample)^2 (every record pair is visited). Seconds per AIREML iteration =
    the GPU fit's operator applications per iteration (GRM product launches, and their
    columns) at that cost; extrapolated."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    ns = min(sample_indiv, c.n)
    nr = int(np.searchsorted(c.indiv, ns))  # records are contiguous per individual, in order
    sub = g.grm_rows(0, ns)
    orc = O.Oracle(c.indiv[:nr], c.times[:nr], c.X[:nr], theta, grm_lt=sub, n=ns)
    rng = np.random.default_rng(0)
    t_k = {}
    for k in (1, 16):
        x = rng.standard_normal((nr, k)).astype(np.float32)
        t0 = time.perf_counter()
        for _ in range(3):
            orc.vx_vc_ref32(x)
        t_k[k] = (time.perf_counter() - t0) / 3
    b = (t_k[16] - t_k[1]) / 15.0
    a = t_k[1] - b
    scale = (c.nrec / nr) ** 2
    t_iter = (a * grm_launches + b * grm_cols) * scale / iters
    return {"sec_per_iter": t_iter, "unit": "s/iteration", "cores": threads, "kind": "port",
            "sample": f"3 + 3 applications of the FP32 record-pair restatement of computeVx_multi_VC on the "
                      f"{nr} records of the first {ns} individuals (k = 1: {t_k[1]:.2f} s, k = 16: "
                      f"{t_k[16]:.2f} s each), a + b k fitted, scaled by (Nrec / {nr})^2 = {scale:.0f} and by "
                      f"the GPU fit's {grm_launches / iters:.1f} operator applications ({grm_cols / iters:.0f} "
                      f"columns) per iteration; extrapolated"}


def aireml_leg(args):
    """AIREML on a dense GRM (N = 40,000, GRM from 80,000 synthetic SNPs built on the
    device, ~320k records, [X | y | 15 Rademacher probes] as one block PCG, start
    0.5 / 0.5 / 0.3 / 0 / 0.3) run to convergence (maxiter 50, tol 1e-3): iterations and
    seconds per AIREML iteration with the GRM product in FP64 and in mixed precision (int8
    digits) from the same start and probes, and the reference operator's CPU cost per
    iteration."""
    import globe_amd as G
    from globe_amd import synth
    g = G.Globe(0)
    rng = np.random.default_rng(args.seed + 11)
    c = synth.simulate_cohort(args.aireml_n, 7.0, rng)
    g.set_cohort(c.indiv, c.times, c.X)
    kept = g.synth_bed(args.aireml_m, seed=args.seed + 11, f_lo=0.05, f_hi=0.5, miss=0.0,
                       min_maf=0.0, max_miss=1.0)
    theta = (0.6, 0.4, 0.4, 0.2, 0.2)
    g.set_varcomp(theta)
    u = np.sqrt(theta[1] / kept) * g.geno_Wa(rng.standard_normal(kept))
    y = synth.simulate_phenotype(c, u, theta, rng)
    t0 = time.perf_counter()
    g.grm_from_geno()
    t_grm = time.perf_counter() - t0
    log(f"aireml: GRM {c.n} x {c.n} from {kept} SNPs built in {t_grm:.2f}s")
    th0 = (0.5, 0.5, 0.3, 0.0, 0.3)
    runs = {}
    for name, prec in (("fp64", G.PREC_FP64), ("mixed", G.PREC_MIXED)):
        if prec == G.PREC_MIXED:
            # the int8 digit tiles of G are built by the first mixed product
            t0 = time.perf_counter()
            g.computeVx_multi(np.zeros((c.nrec, 1)), multGRM_uhat=True, precision=prec)
            t_dig = time.perf_counter() - t0
            log(f"aireml: int8 digit tiles of G built in {t_dig:.2f}s")
        mark(g)
        fit = g.aireml_fit(y, th0, max_iter=args.aireml_iters, tol=1e-3, num_mc=15, probes="rademacher",
                           seed=args.seed, pcg_tol=1e-5, precision=prec)
        mark(g)
        # the GRM products' HIP-event times from a second, identical fit (timing every
        # launch takes the host's read of each iteration's events, so it is not the timed one)
        g.profile(True)
        g.aireml_fit(y, th0, max_iter=args.aireml_iters, tol=1e-3, num_mc=15, probes="rademacher",
                     seed=args.seed, pcg_tol=1e-5, precision=prec)
        ks = g.kernel_stats(3 if prec == G.PREC_MIXED else 2)
        g.profile(False)
        if name == "fp64":
            grm_launches = ks["launches"]
            grm_cols = ks["flops"] * ks["launches"] / (2.0 * c.n * c.n)
        runs[name] = {"sec_per_iter": fit["sec_per_iter"], "iters": fit["iters"],
                      "theta": list(map(float, fit["theta"])), "grm_passes": ks["launches"],
                      "grm_pass_ms": ks["ms_avg"],
                      "grm_pass_GBps": ks["bytes"] / (ks["ms_avg"] * 1e-3) / 1e9 if ks["ms_avg"] else None}
        log(f"aireml {name}: {fit['iters']} iterations, {fit['sec_per_iter']:.3f} s/iter, GRM pass "
            f"{ks['ms_avg']:.3f} ms x {ks['launches']}, theta {np.round(fit['theta'], 5)}")
        # the REML log-likelihood at the FP64 fit's components: log|V| by Lanczos
        # quadrature over 64 Rademacher probe solves (one block PCG), same probes for both
        if name == "fp64":
            th_ll = tuple(fit["theta"])
        g.set_varcomp(th_ll)
        ll = g.reml_loglik(y, probes="rademacher", nprobe=64, seed=args.seed, tol=1e-6, precision=prec)
        runs[name]["loglik"] = {k: ll[k] for k in ("loglik", "logdet_v", "logdet_v_se", "logdet_q", "ypy",
                                                   "max_steps", "seconds")}
        log(f"aireml {name}: REML log-likelihood {ll['loglik']:.6f} (log|V| {ll['logdet_v']:.4f} +/- "
            f"{ll['logdet_v_se']:.3g}, {ll['max_steps']} Lanczos steps) in {ll['seconds']:.3f}s")
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log("aireml: CPU baseline (record-pair operator) ...")
        cpu = aireml_cpu_baseline(g, c, theta, grm_launches, grm_cols, runs["fp64"]["iters"],
                                  args.aireml_cpu_indiv, threads)
        log(f"aireml: CPU baseline {cpu['sec_per_iter']:.0f} s per iteration ({cpu['kind']})")
    t64, tm = np.array(runs["fp64"]["theta"]), np.array(runs["mixed"]["theta"])
    out = {"n": c.n, "m_snps": int(kept), "nrec": c.nrec, "probes": 15, "grm_build_s": t_grm,
           "iters": runs["fp64"]["iters"], "sec_per_iter": runs["fp64"]["sec_per_iter"],
           "theta": runs["fp64"]["theta"], "precision": "fp64 (FP32 GRM on the FP64 matrix cores)",
           "mixed": dict(runs["mixed"], precision="mixed (int8 digits of G and S0, relaxed 4/3/2)",
                         digit_build_s=t_dig),
           "mixed_vs_fp64_speedup": runs["fp64"]["sec_per_iter"] / runs["mixed"]["sec_per_iter"],
           "mixed_theta_max_rel_diff": float(np.max(np.abs(tm - t64) / np.abs(t64))),
           "fp64_grm_pass_ms": runs["fp64"]["grm_pass_ms"],
           "loglik": dict(runs["fp64"]["loglik"], probes=64, at="the FP64 fit's components",
                          mixed=runs["mixed"]["loglik"]),
           "cpu_baseline": cpu,
           "note": f"the fit from the same start and probes, maxiter {args.aireml_iters}, tol 1e-3"}
    g.close()
    return out


def single_rhs_leg(args):
    """The single-RHS runPCGsolver --geno workload at N = 50,000, M = 100,000 SNPs (1%
    missing, QC minMAF 0.01 / maxMissingRate 0.05), 1 + Poisson(7) records (~400k): the
    solve of y in FP64 (7-digit int8 contraction), in the FP64 vector-ALU form and in mixed
    precision (4 digits, relaxed), tol 1e-6 -- solves/s, iterations, the packed-genotype
    passes' time and HBM rate, and each solution's error against the FP64 one."""
    import globe_amd as G
    g = G.Globe(0)
    c, y, kept = build_cohort(g, args.cfg1_n, args.cfg1_m, args.lam, args.seed + 17)
    nrec = c.nrec
    b1, x1 = g.alloc(nrec * 8), g.alloc(nrec * 8)
    b1.upload(np.ascontiguousarray(y))
    out = {"n": c.n, "m_snps": args.cfg1_m, "m_kept": int(kept), "nrec": nrec, "tol": args.tol}
    ref = None
    for name, prec in (("fp64", G.PREC_FP64), ("fp64_valu", G.PREC_FP64_VALU), ("mixed", G.PREC_MIXED)):
        g.pcg_dev(b1, x1, 1, tol=args.tol, rule=G.RULE_PER_COLUMN, precision=prec)  # warm (tuning)
        reps, t_sum = 5, 0.0
        for _ in range(reps):
            st = g.pcg_dev(b1, x1, 1, tol=args.tol, rule=G.RULE_PER_COLUMN, precision=prec)
            t_sum += st["seconds"]
        # pass times from one more, profiled solve (not timed: profiling reads every
        # iteration's events on the host)
        g.profile(True)
        g.pcg_dev(b1, x1, 1, tol=args.tol, rule=G.RULE_PER_COLUMN, precision=prec)
        p = [g.kernel_stats(0), g.kernel_stats(1)]
        g.profile(False)
        x = x1.download(np.zeros(nrec))
        r = {"solves_per_s": reps / t_sum, "seconds": t_sum / reps, "iters": st["iters"]}
        if prec != G.PREC_FP64_VALU:
            r.update({"pass1_ms": p[0]["ms_avg"], "pass2_ms": p[1]["ms_avg"],
                      "pass_GBps": [q["bytes"] / (q["ms_avg"] * 1e-3) / 1e9 if q["launches"] else None for q in p],
                      "hbm_frac": [q["bytes"] / (q["ms_avg"] * 1e-3) / 1e9 / PEAK_HBM_GBPS if q["launches"] else None
                                   for q in p]})
        if ref is None:
            ref = x
        else:
            r["rel_err_vs_fp64"] = float(np.linalg.norm(x - ref) / np.linalg.norm(ref))
        out[name] = r
        log(f"single-RHS {name}: {st['iters']} iters, {r['seconds'] * 1e3:.1f} ms, {r['solves_per_s']:.1f} solves/s"
            + (f", passes {p[0]['ms_avg']:.3f} / {p[1]['ms_avg']:.3f} ms" if "pass1_ms" in r else ""))
    g.close()
    return out


def grm_leg(args):
    """The --grm path at the genotype workload's cohort size (N = 200,000, ~1.6M
    records): G = W W^T / M built on the device from args.grm_snps synthetic SNPs into
    the HBM-resident tile-major FP32 lower triangle (80 GB), then PCG on the GRM product:
    the FP64 1-RHS solve of y and 32-RHS block solves ([y | 31 Rademacher columns]) in
    FP64 and mixed precision (int8 digit tiles of G, 160 GB, built by the first mixed
    product). Runs after the genotype context is closed."""
    import globe_amd as G
    from globe_amd import synth
    g = G.Globe(0)
    rng = np.random.default_rng(args.seed + 13)
    c = synth.simulate_cohort(args.n, args.lam, rng)
    g.set_cohort(c.indiv, c.times, c.X)
    kept = g.synth_bed(args.grm_snps, seed=args.seed + 13, f_lo=0.05, f_hi=0.5, miss=0.0,
                       min_maf=0.0, max_miss=1.0)
    theta = (0.6, 0.4, 0.4, 0.2, 0.2)
    g.set_varcomp(theta)
    u = np.sqrt(theta[1] / kept) * g.geno_Wa(rng.standard_normal(kept))
    y = synth.simulate_phenotype(c, u, theta, rng)
    t0 = time.perf_counter()
    g.grm_from_geno()
    t_build = time.perf_counter() - t0
    log(f"grm: G {c.n} x {c.n} from {kept} SNPs built in {t_build:.2f}s")
    nrec, k = c.nrec, args.k
    B = np.asfortranarray(np.column_stack([y, rng.choice([-1.0, 1.0], size=(nrec, k - 1))]))
    bbuf, xbuf = g.alloc(B.nbytes), g.alloc(B.nbytes)
    bbuf.upload(B)
    out = {"n": c.n, "nrec": nrec, "m_snps": int(kept), "build_s": t_build, "rhs": k, "tol": args.tol,
           "grm_bytes": ((c.n + 63) // 64) * ((c.n + 63) // 64 + 1) // 2 * 4096 * 4}
    runs = [("fp64_1rhs", 1, G.PREC_FP64, 2), ("fp64_block", k, G.PREC_FP64, 2), ("mixed_block", k, G.PREC_MIXED, 3)]
    x0 = None
    for name, kk, prec, which in runs:
        try:
            g.pcg_dev(bbuf, xbuf, kk, tol=args.tol, rule=G.RULE_PER_COLUMN, precision=prec)  # warm
            st = g.pcg_dev(bbuf, xbuf, kk, tol=args.tol, rule=G.RULE_PER_COLUMN, precision=prec)
            g.profile(True)
            g.pcg_dev(bbuf, xbuf, kk, tol=args.tol, rule=G.RULE_PER_COLUMN, precision=prec)
            ks = g.kernel_stats(which)
            g.profile(False)
        except G.GlobeError as e:
            out[name] = {"error": str(e)}
            log(f"grm {name}: {e}")
            continue
        conv = int(np.sum(st["col_ratio"][:kk] < args.tol))
        X = xbuf.download(np.zeros((nrec, k), order="F"))
        r = {"solves_per_s": conv / st["seconds"], "seconds": st["seconds"], "iters": st["iters"],
             "grm_pass_ms": ks["ms_avg"], "grm_passes": ks["launches"],
             "grm_pass_GBps": ks["bytes"] / (ks["ms_avg"] * 1e-3) / 1e9 if ks["ms_avg"] else None}
        if name == "fp64_1rhs":
            x0 = X[:, 0].copy()
        else:
            r["y_rel_err_vs_fp64_1rhs"] = float(np.linalg.norm(X[:, 0] - x0) / np.linalg.norm(x0))
        out[name] = r
        log(f"grm {name}: {st['iters']} iters, {st['seconds']:.3f}s, {r['solves_per_s']:.2f} solves/s, "
            f"GRM pass {ks['ms_avg']:.2f} ms x {ks['launches']}")
    g.close()
    return out


def ingest_leg(g, c, args):
    """Genotype ingest from a PLINK file set and from a host buffer into a second
    context: the first args.ingest_snps kept rows of the resident synthetic cohort,
    read back as PLINK codes and written to a .bed/.bim/.fam under the temp folder
    (page cache warm, so the rate is host memory -> HBM, not the disk)."""
    import tempfile
    import globe_amd as G
    n = c.n
    nb = (n + 3) // 4
    rows = np.ascontiguousarray(g.geno_rows(0, min(args.ingest_snps, g.m))[:, :nb])
    m = rows.shape[0]
    tmp = tempfile.mkdtemp(prefix="globe_ingest_")
    prefix = os.path.join(tmp, "cohort")
    t0 = time.perf_counter()
    with open(prefix + ".bed", "wb") as fh:
        fh.write(bytes([0x6C, 0x1B, 0x01]))
        fh.write(rows.tobytes())
    with open(prefix + ".bim", "w") as fh:
        fh.writelines(f"1\trs{s + 1}\t0\t{s + 1}\tA\tG\n" for s in range(m))
    with open(prefix + ".fam", "w") as fh:
        fh.writelines(f"{i + 1} {i + 1} 0 0 0 -9\n" for i in range(n))
    log(f"ingest: wrote {rows.nbytes / 1e9:.2f} GB .bed ({m} SNPs x {n}) in {time.perf_counter() - t0:.1f}s")
    h = G.Globe(0)
    h.set_cohort(c.indiv, c.times, c.X)
    out = {"snps": m, "individuals": n, "bed_bytes": int(rows.nbytes) + 3}
    try:
        h.load_bed_file(prefix, 0.01, 0.05)  # warm-up: page cache, pinned pools
        for name, run in (("file", lambda: h.load_bed_file(prefix, 0.01, 0.05)),
                          ("buffer", lambda: h.prepGenotypeInput(rows, 0.01, 0.05))):
            t1 = time.perf_counter()
            kept = run()
            dt = time.perf_counter() - t1
            st = h.ingest_stats()
            out[name] = {"seconds": dt, "GBps": rows.nbytes / dt / 1e9, "kept": kept,
                         "host_fill_s": st["fill_s"], "host_wait_s": st["wait_s"],
                         "chunks": st["chunks"], "host_threads": st["threads"]}
            log(f"ingest from {name}: {rows.nbytes / dt / 1e9:.1f} GB/s ({dt:.3f}s, host fill "
                f"{st['fill_s']:.3f}s, waiting {st['wait_s']:.3f}s, {st['threads']} threads)")
    finally:
        h.close()
        for ext in (".bed", ".bim", ".fam"):
            os.remove(prefix + ext)
        os.rmdir(tmp)
    out["note"] = ("pinned triple-buffered chunks of 256 MiB, parallel pread/memcpy, async copies; "
                   "QC counts, keep rule and compaction on the device; ends with both layouts built")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--m", type=int, default=500_000)
    ap.add_argument("--lam", type=float, default=7.0)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=20261206)
    ap.add_argument("--cpu-sample-snps", type=int, default=20_000)
    ap.add_argument("--cpu-ref-snps", type=int, nargs="+", default=[1_000, 4_000, 50_000],
                    help="SNP samples on which the reference runPCGsolver is timed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mark", action="store_true",
                    help="delimit the serving replays and AI-REML fits with 16-byte device copies")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests replayed through the serving loop (0: skip)")
    ap.add_argument("--serve-hbm-frac", type=float, default=0.05,
                    help="share of the free HBM given to the serving loop's record pool")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-slots", type=int, default=32)
    ap.add_argument("--ingest-snps", type=int, default=40_000,
                    help="SNPs of the .bed written and ingested (0: skip)")
    ap.add_argument("--aireml-iters", type=int, default=50, help="AIREML maxiter (0: skip)")
    ap.add_argument("--aireml-cpu-indiv", type=int, default=10_000,
                    help="individuals whose records time the CPU record-pair operator")
    ap.add_argument("--aireml-n", type=int, default=40_000)
    ap.add_argument("--aireml-m", type=int, default=80_000)
    ap.add_argument("--cfg1-n", type=int, default=50_000, help="single-RHS workload individuals (0: skip)")
    ap.add_argument("--cfg1-m", type=int, default=100_000, help="single-RHS workload SNPs")
    ap.add_argument("--grm-snps", type=int, default=20_000,
                    help="SNPs of the device-built GRM of the N-individual --grm leg (0: skip)")
    ap.add_argument("--calib", action="store_true",
                    help="first run 3 plain 1 GiB device copies (counter calibration)")
    args = ap.parse_args()
    global MARK
    MARK = args.mark
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; SNP-sharded runs over several "
                                   "GPUs are not implemented", "n_gpus": args.gpus}))
        sys.exit(2)

    import globe_amd as G
    g = G.Globe(0)
    if args.calib:
        # reference dispatches for the HBM byte counters: 3 plain copies of 1 GiB
        a, b = g.alloc(1 << 30), g.alloc(1 << 30)
        for _ in range(3):
            g._check(g.L.globe_dev_copy(g.h, b.ptr, a.ptr, 1 << 30))
        a.free()
        b.free()
    c, y, kept = build_cohort(g, args.n, args.m, args.lam, args.seed)
    nrec, k = c.nrec, args.k
    rng = np.random.default_rng(args.seed + 1)
    B = np.asfortranarray(np.column_stack([y, rng.choice([-1.0, 1.0], size=(nrec, k - 1))]))
    bbuf = g.alloc(B.nbytes)
    xbuf = g.alloc(B.nbytes)
    bbuf.upload(B)
    b1 = g.alloc(nrec * 8)
    x1 = g.alloc(nrec * 8)
    b1.upload(np.ascontiguousarray(y))

    # FP64 single-RHS solve of y (the comparison point)
    log("FP64 1-RHS solve ...")
    # untimed first solve: the contraction tuner times its candidates on this call
    g.pcg_dev(b1, x1, 1, tol=args.tol, rule=G.RULE_PER_COLUMN, precision=G.PREC_FP64)
    g.profile(True)
    st64 = g.pcg_dev(b1, x1, 1, tol=args.tol, rule=G.RULE_PER_COLUMN, precision=G.PREC_FP64)
    p64 = [g.kernel_stats(0), g.kernel_stats(1)]
    x64 = x1.download(np.zeros(nrec))
    log(f"FP64 1-RHS: {st64['iters']} iters, {st64['seconds']:.3f}s, pass1 {p64[0]['ms_avg']:.3f} ms, "
        f"pass2 {p64[1]['ms_avg']:.3f} ms")
    # the plain FP64 form of the same solve: FP64 FMAs over the dosages on the vector ALUs
    g.profile(False)
    stv = g.pcg_dev(b1, x1, 1, tol=args.tol, rule=G.RULE_PER_COLUMN, precision=G.PREC_FP64_VALU)
    xv = x1.download(np.zeros(nrec))
    valu = {"solves_per_s": 1.0 / stv["seconds"], "iters": stv["iters"], "seconds": stv["seconds"],
            "rel_err_vs_fp64": float(np.linalg.norm(xv - x64) / np.linalg.norm(x64))}
    log(f"FP64 VALU 1-RHS: {stv['iters']} iters, {stv['seconds']:.3f}s")

    for w in range(args.warmup):
        log(f"warmup {w}")
        g.pcg_dev(bbuf, xbuf, k, tol=args.tol, rule=G.RULE_PER_COLUMN, precision=G.PREC_MIXED)
    g.profile(True)
    g.synchronize()
    t0 = time.perf_counter()
    stats = []
    for s in range(args.steps):
        stats.append(g.pcg_dev(bbuf, xbuf, k, tol=args.tol, rule=G.RULE_PER_COLUMN,
                               precision=G.PREC_MIXED))
        log(f"step {s}: {stats[-1]['iters']} iters, {stats[-1]['seconds']:.3f}s")
    g.synchronize()
    wall = time.perf_counter() - t0
    pm = [g.kernel_stats(0), g.kernel_stats(1)]
    g.profile(False)
    X = xbuf.download(np.zeros((nrec, k), order="F"))
    conv = sum(int(np.sum(st["col_ratio"] < args.tol)) for st in stats)
    err_y = float(np.linalg.norm(X[:, 0] - x64) / np.linalg.norm(x64))
    value = conv / wall

    # dominant kernel: the mixed contraction (both passes share the kernel); roofline
    # against int8 MFMA peak with algorithmic flops
    tot_ms = pm[0]["ms_total"] + pm[1]["ms_total"]
    launches = pm[0]["launches"] + pm[1]["launches"]
    flops = pm[0]["flops"]
    achieved_tflops = flops / (tot_ms / launches * 1e-3) / 1e12 if launches else 0.0
    issued = (pm[0]["issued"] * pm[0]["launches"] + pm[1]["issued"] * pm[1]["launches"])
    issued_tops = issued / (tot_ms * 1e-3) / 1e12 if tot_ms else 0.0
    # 1-RHS FP64 pass HBM roofline (algorithmic packed bytes)
    gbps64 = [p["bytes"] / (p["ms_avg"] * 1e-3) / 1e9 if p["launches"] else 0.0 for p in p64]

    serve = None
    if args.serve_requests > 0:
        serve = serving_leg(g, c, kept, args)
    ingest = ingest_leg(g, c, args) if args.ingest_snps > 0 else None
    aireml = aireml_leg(args) if args.aireml_iters > 0 else None
    traffic, traffic_src = measured_traffic("wide")
    if traffic_src:
        traffic_src = f"profiles/k_geno_mfma_n200k_m500k.md ({traffic_src})"
    traffic64, traffic64_src = measured_traffic("thin")
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log("CPU baseline: the reference runPCGsolver ...")
        cpu = cpu_baseline_reference(g, c, y, kept, st64["iters"],
                                     sorted({min(s_, kept) for s_ in args.cpu_ref_snps}),
                                     threads, (0.6, 0.4, 0.4, 0.2, 0.2))
        log("CPU baseline: the oracle's FP32 restatement ...")
        port = cpu_baseline(g, c, kept, st64["iters"], min(args.cpu_sample_snps, kept), threads)
        if cpu is None:
            cpu = port
        else:
            cpu["port"] = port
        log(f"CPU baseline: {cpu['value']:.5f} solves/s ({cpu['kind']})")
    out = {
        "metric": "PCG solves/s at N=200k,M=500k (mixed vs FP64); AIREML iter s; p50/p99",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8-mfma(4 digits)+f64",
        "data": "synthetic (genotypes generated in HBM, LGMM phenotype)",
        "config": {"workload": "runPCGsolver --geno block PCG", "n": args.n, "m_snps": args.m,
                   "m_kept": kept, "nrec": nrec, "rhs": k, "tol": args.tol,
                   "precision": "mixed (int8 MFMA, 4 digits) vs FP64 (7 digits)"},
        "fp64_1rhs": {"solves_per_s": 1.0 / st64["seconds"], "iters": st64["iters"],
                      "seconds": st64["seconds"],
                      "pass1_ms": p64[0]["ms_avg"], "pass2_ms": p64[1]["ms_avg"],
                      "pass1_GBps": gbps64[0], "pass2_GBps": gbps64[1],
                      "hbm_frac": [x / PEAK_HBM_GBPS for x in gbps64],
                      "traffic": traffic64, "traffic_source": traffic64_src},
        "fp64_valu_1rhs": valu,
        "mixed_vs_fp64_speedup": value * st64["seconds"],
        "mixed_vs_fp64_valu_speedup": value * stv["seconds"],
        "mixed_solution_rel_err_y": err_y,
        "mixed_iters": [st["iters"] for st in stats],
        "roofline": {"bound": "mfma", "achieved": achieved_tflops, "peak": PEAK_INT8_TOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_INT8_TOPS, "traffic": traffic,
                     "traffic_source": traffic_src, "algorithmic_bytes": pm[0]["bytes"],
                     "kernel": "k_geno_wide (both passes)", "ms_per_launch": tot_ms / max(launches, 1),
                     "issued_int8_TOPs": issued_tops, "issued_frac": issued_tops / PEAK_INT8_TOPS,
                     "mean_digits": (issued / launches) / flops if launches and flops else None},
        "cpu_baseline": cpu,
        "serving": serve,
        "aireml": aireml,
        "ingest": ingest,
    }
    g.close()
    if args.cfg1_n > 0:
        out["single_rhs"] = single_rhs_leg(args)
    if args.grm_snps > 0:
        out["grm"] = grm_leg(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
