"""Write the golden fixtures of tests/golden/ from the compiled reference.

Runs oracle/_ref/runPCGsolver (built by `make -C oracle ref` from the reference's
own sources) on small synthetic --geno and --grm cases, twice each (the reference
starts PCG from an uninitialised matrix, PCGsolver.cpp:265, so both runs must
agree), and stores inputs plus the reference's _beta/_uhat/_bhat vectors as .npz.
Run in the build container only; the fixtures are data.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from globe_amd import synth  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "runPCGsolver")
OUT = os.path.dirname(os.path.abspath(__file__))

CASES = [
    # name, kind, n, m, lam, seed, miss, theta, (f_lo, f_hi), (minMAF, maxMissingRate)
    ("geno_n203", "geno", 203, 700, 3.0, 11, 0.01, (0.6, 0.4, 0.4, 0.2, 0.2), (0.05, 0.5), (0.01, 0.2)),
    ("geno_n150_odd", "geno", 150, 400, 2.0, 12, 0.02, (0.5, 0.3, 0.3, 0.05, 0.1), (0.05, 0.5), (0.01, 0.2)),
    # QC stress: frequencies and missingness straddle the thresholds
    ("geno_qc_n257", "geno", 257, 900, 2.0, 14, 0.04, (0.6, 0.4, 0.4, 0.2, 0.2), (0.0, 0.12), (0.05, 0.045)),
    ("grm_n180", "grm", 180, 500, 3.0, 13, 0.0, (0.6, 0.4, 0.4, 0.2, 0.2), (0.05, 0.5), None),
]


def run_ref(paths, kind, theta, out_prefix, extra=()):
    cmd = [REF, "--input", paths["input"], "--output", out_prefix,
           "--sig2e", repr(theta[0]), "--sig2g", repr(theta[1]), "--sig2b0", repr(theta[2]),
           "--sigb01", repr(theta[3]), "--sig2b1", repr(theta[4]), *extra]
    cmd += ["--geno", paths["geno"]] if kind == "geno" else ["--grm", paths["grm"]]
    r = subprocess.run(cmd, capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
    if r.returncode != 0:
        raise RuntimeError(r.stdout + r.stderr)
    iters = [int(l.split()[-1]) for l in r.stdout.splitlines() if "Converged at iteration" in l]
    kept = [int(l.split("=")[1].split("/")[0]) for l in r.stdout.splitlines() if "Filtered SNPs" in l]
    out = {s: synth.read_vector_bin(f"{out_prefix}_{s}.bin") for s in ("beta", "uhat", "bhat")}
    return out, iters, kept


def main():
    if not os.path.exists(REF):
        raise SystemExit(f"{REF} missing: run `make -C oracle ref` first")
    for name, kind, n, m, lam, seed, miss, theta, (f_lo, f_hi), qc in CASES:
        extra = ["--debug"]
        if kind == "geno":
            d = synth.make_geno_case(n, m, lam, seed, f_lo=f_lo, f_hi=f_hi, miss=miss, theta=theta,
                                     min_maf=qc[0], max_miss=qc[1])
            extra += ["--minMAF", repr(qc[0]), "--maxMissingRate", repr(qc[1])]
        else:
            d = synth.make_grm_case(n, m, lam, seed, f_lo=f_lo, f_hi=f_hi, theta=theta)
        with tempfile.TemporaryDirectory() as tmp:
            paths = synth.write_case_files(d, tmp, name)
            r1, it1, kept1 = run_ref(paths, kind, theta, os.path.join(tmp, "o1"), extra)
            r2, it2, kept2 = run_ref(paths, kind, theta, os.path.join(tmp, "o2"), extra)
        for s in r1:
            assert np.array_equal(r1[s], r2[s]), f"{name}: reference runs disagree on {s}"
        c = d.cohort
        arrs = dict(indiv=c.indiv, times=c.times, covars=c.covars, y=d.y, theta=np.array(theta),
                    ref_beta=r1["beta"], ref_uhat=r1["uhat"], ref_bhat=r1["bhat"],
                    ref_iters=np.array(it1))
        if kind == "geno":
            assert kept1 and kept1 == kept2
            arrs.update(bed_rows=d.bed_rows, min_maf=np.float32(qc[0]), max_miss=np.float32(qc[1]),
                        ref_kept=np.int64(kept1[0]))
        else:
            arrs.update(grm_lt=d.grm_lt)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrs)
        print(name, "iters", it1, "kept", kept1, "beta", r1["beta"])


if __name__ == "__main__":
    main()
