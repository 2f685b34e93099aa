"""The drop-in CLIs: same flags, files and exit behaviour as the reference.

Flag and file errors exit with status 1 before any device work (CPU tests); the
solves are GPU tests comparing the written _beta/_uhat/_bhat.bin files with the
compiled reference's (golden fixtures) and with the FP64 oracle.
"""
import os
import subprocess

import numpy as np
import pytest

from globe_amd import synth
from helpers import golden_oracle, load_golden, rel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PCG = os.path.join(ROOT, "globe_amd", "bin", "runPCGsolver")
AIREML = os.path.join(ROOT, "globe_amd", "bin", "runAIREML")
MAKEGRM = os.path.join(ROOT, "globe_amd", "bin", "makeGRM")
SIG = ["--sig2e", "0.6", "--sig2g", "0.4", "--sig2b0", "0.4", "--sigb01", "0.2", "--sig2b1", "0.2"]


def run(args, **kw):
    return subprocess.run(args, capture_output=True, text=True, timeout=300, **kw)


@pytest.fixture(scope="module")
def case_files(tmp_path_factory):
    d = load_golden("geno_n203")
    tmp = str(tmp_path_factory.mktemp("cli"))
    c = d["cohort"]
    sd = synth.SynthData(cohort=c, y=d["y"], bed_rows=d["bed_rows"], theta=tuple(d["theta"]))
    paths = synth.write_case_files(sd, tmp, "geno")
    g = load_golden("grm_n180")
    gd = synth.SynthData(cohort=g["cohort"], y=g["y"], grm_lt=g["grm_lt"], theta=tuple(g["theta"]))
    paths.update({("grm_" + k): v for k, v in synth.write_case_files(gd, tmp, "grm").items()})
    paths["dir"] = tmp
    return paths


def test_usage_and_flag_errors(case_files):
    p = case_files
    assert run([PCG]).returncode == 1
    r = run([PCG, "--input", p["input"], "--output", "x", "--geno", p["geno"]])
    assert r.returncode == 1 and "Missing required flags" in r.stderr
    r = run([PCG, "--input", p["input"], "--output", "x", *SIG, "--geno", p["geno"], "--grm", p["grm_grm"]])
    assert r.returncode == 1 and "exactly one of" in r.stderr
    r = run([PCG, "--input", p["input"], "--output", "x", *SIG])
    assert r.returncode == 1 and "exactly one of" in r.stderr
    r = run([PCG, "--input", p["input"], "--output", "x", *SIG, "--geno", p["geno"], "--bogus"])
    assert r.returncode == 1 and "Unknown flag: --bogus" in r.stderr
    r = run([PCG, "--input", p["input"], "--output", "x", *SIG, "--geno", p["geno"], "--minMAF", "0.7"])
    assert r.returncode == 1 and "Invalid MAF" in r.stderr
    r = run([PCG, "--input", p["input"], "--output", "x", *SIG, "--geno", p["geno"], "--sig2e"])
    assert r.returncode == 1 and "expects an argument" in r.stderr
    assert run([AIREML, "a", "b"]).returncode == 1


def test_file_errors(case_files, tmp_path):
    p = case_files
    bad = str(tmp_path / "bad")
    with open(bad + ".bed", "wb") as fh:
        fh.write(bytes([0x6C, 0x1B, 0x00]) + bytes(100))
    open(bad + ".bim", "w").write("1 rs1 0 1 A G\n")
    r = run([PCG, "--input", p["input"], "--output", "x", *SIG, "--geno", bad])
    assert r.returncode == 1 and "individual-major" in r.stderr
    with open(bad + ".bed", "wb") as fh:
        fh.write(bytes([0x6C, 0x1C, 0x01]))
    r = run([PCG, "--input", p["input"], "--output", "x", *SIG, "--geno", bad])
    assert r.returncode == 1 and "magic" in r.stderr
    with open(bad + ".grm.bin", "wb") as fh:
        fh.write(bytes(16))
    r = run([PCG, "--input", p["grm_input"], "--output", "x", *SIG, "--grm", bad])
    assert r.returncode == 1 and "Error reading" in r.stderr
    r = run([PCG, "--input", str(tmp_path / "missing.txt"), "--output", "x", *SIG, "--grm", bad])
    assert r.returncode == 1
    txt = str(tmp_path / "nohdr.txt")
    open(txt, "w").write("ID PHENO TIMES\n1 0.5 1.0\n")
    r = run([PCG, "--input", txt, "--output", "x", *SIG, "--grm", bad])
    assert r.returncode == 1 and "Missing required columns" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["geno", "grm"])
def test_runpcgsolver_matches_reference_and_oracle(kind, case_files):
    p = case_files
    name = "geno_n203" if kind == "geno" else "grm_n180"
    d = load_golden(name)
    out = os.path.join(p["dir"], "out_" + kind)
    inp = p["input"] if kind == "geno" else p["grm_input"]
    src = ["--geno", p["geno"]] if kind == "geno" else ["--grm", p["grm_grm"]]
    print("[case] runPCGsolver", kind, flush=True)
    r = run([PCG, "--input", inp, "--output", out, *SIG, *src, "--debug"])
    assert r.returncode == 0, r.stderr + r.stdout
    assert "[main] Done." in r.stdout
    res = {s: synth.read_vector_bin(f"{out}_{s}.bin") for s in ("beta", "uhat", "bhat")}
    assert res["beta"].shape == (4,) and res["uhat"].shape == (d["cohort"].n,)
    for s in res:
        assert rel(res[s], d["ref_" + s]) < 5e-4, s
    orc = golden_oracle(d)
    ro = orc.gls_blup(d["y"], tol=1e-5, stagnation=True)
    for s in ("beta", "uhat", "bhat"):
        # FP32 output files against the FP64 oracle run with the same control flow
        assert rel(res[s], ro[s]) < 1e-6, s


@pytest.mark.gpu
def test_runaireml_runs_and_logs(case_files):
    p = case_files
    out = os.path.join(p["dir"], "aireml")
    print("[case] runAIREML", flush=True)
    r = run([AIREML, p["grm_input"], p["grm_grm"], out, "--maxiter", "3", "--numMC", "4",
             "--sig2e", "0.5", "--sig2g", "0.5", "--sig2b0", "0.3", "--sigb01", "0.0", "--sig2b1", "0.3"])
    assert r.returncode == 0, r.stderr + r.stdout
    log = open(out + ".log").read()
    assert "Iteration 1 took" in log and "Final variance components" in log
    assert log.startswith("Command:")
    assert not os.path.exists(out + "_beta.bin")  # the reference writes no solutions


@pytest.mark.gpu
def test_runaireml_write_solutions(case_files):
    """--writeSolutions: GLS / BLUP at the fitted components into _beta/_uhat/_bhat.bin
    (the reference's commented-out saveVector calls, varCompEst.cpp:1493-1495)."""
    p = case_files
    out = os.path.join(p["dir"], "aireml_sol")
    print("[case] runAIREML --writeSolutions", flush=True)
    r = run([AIREML, p["grm_input"], p["grm_grm"], out, "--maxiter", "2", "--numMC", "4", "--writeSolutions",
             "--sig2e", "0.5", "--sig2g", "0.5", "--sig2b0", "0.3", "--sigb01", "0.0", "--sig2b1", "0.3"])
    assert r.returncode == 0, r.stderr + r.stdout
    log = open(out + ".log").read()
    theta = []
    for name in ("sig2e", "sig2g", "sig2b0", "sigb01", "sig2b1"):
        line = next(ln for ln in log.splitlines() if ln.strip().startswith(name + " ="))
        theta.append(float(line.split("=")[1].split("+/-")[0]))
    d = load_golden("grm_n180")
    res = {s: synth.read_vector_bin(f"{out}_{s}.bin") for s in ("beta", "uhat", "bhat")}
    assert res["beta"].shape == (4,) and res["bhat"].shape == (2 * d["cohort"].n,)
    # the log prints 6 significant digits of the components
    ro = golden_oracle(d, theta=tuple(theta)).gls_blup(d["y"], tol=1e-8, x0_mode=1, rule=1)
    for s in res:
        assert rel(res[s], ro[s]) < 1e-4, s


@pytest.mark.gpu
def test_runaireml_loglik_exact(case_files):
    """--loglik exact: the REML log-likelihood at the fitted components equals the dense
    closed form at the components the log prints (6 significant digits)."""
    p = case_files
    out = os.path.join(p["dir"], "aireml_ll")
    print("[case] runAIREML --loglik exact", flush=True)
    r = run([AIREML, p["grm_input"], p["grm_grm"], out, "--maxiter", "2", "--numMC", "4", "--loglik", "exact",
             "--pcgTol", "1e-10", "--sig2e", "0.5", "--sig2g", "0.5", "--sig2b0", "0.3", "--sigb01", "0.0",
             "--sig2b1", "0.3"])
    assert r.returncode == 0, r.stderr + r.stdout
    log = open(out + ".log").read()
    theta = []
    for name in ("sig2e", "sig2g", "sig2b0", "sigb01", "sig2b1"):
        line = next(ln for ln in log.splitlines() if ln.strip().startswith(name + " ="))
        theta.append(float(line.split("=")[1].split("+/-")[0]))
    line = next(ln for ln in log.splitlines() if ln.startswith("REML log-likelihood"))
    ll = float(line.split(":")[1].split("(")[0])
    d = load_golden("grm_n180")
    c = d["cohort"]
    G = synth.grm_dense(d["grm_lt"], c.n)
    from helpers import dense_V
    V, _ = dense_V(c, G, theta)
    import oracle as O
    ref = O.reml_loglik_dense(V, c.X, d["y"])[0]
    # the printed components carry 6 digits; the log-likelihood moves by far less than
    # 1e-6 of itself over that rounding near the fitted point
    assert abs(ll - ref) < 1e-6 * abs(ref), (ll, ref)
    r = run([AIREML, p["grm_input"], p["grm_grm"], out, "--maxiter", "1", "--loglik", "1"])
    assert r.returncode == 1 and "--loglik" in r.stderr


def test_makegrm_flag_and_file_errors(case_files, tmp_path):
    p = case_files
    r = run([MAKEGRM])
    assert r.returncode == 1 and "Usage" in r.stderr
    r = run([MAKEGRM, "--geno", p["geno"], "--output", str(tmp_path / "g"), "--bogus"])
    assert r.returncode == 1 and "Unknown flag: --bogus" in r.stderr
    r = run([MAKEGRM, "--geno", p["geno"], "--output", str(tmp_path / "g"), "--minMAF", "0.7"])
    assert r.returncode == 1 and "Invalid MAF" in r.stderr
    r = run([MAKEGRM, "--geno", str(tmp_path / "nothing"), "--output", str(tmp_path / "g")])
    assert r.returncode == 1 and ".fam" in r.stderr
    assert not os.path.exists(str(tmp_path / "g.grm.bin"))


@pytest.mark.gpu
def test_makegrm_writes_wwt_and_round_trips(case_files):
    """makeGRM writes G = W W^T / M over the QC-kept SNPs as a .grm.bin; runPCGsolver
    --grm on that file reproduces runPCGsolver --geno."""
    p = case_files
    d = load_golden("geno_n203")
    c = d["cohort"]
    out = os.path.join(p["dir"], "made")
    print("[case] makeGRM", flush=True)
    r = run([MAKEGRM, "--geno", p["geno"], "--output", out, "--minMAF", str(float(d["min_maf"])),
             "--maxMissingRate", str(float(d["max_miss"]))])
    assert r.returncode == 0, r.stderr + r.stdout
    lt = np.fromfile(out + ".grm.bin", dtype=np.float32)
    assert lt.shape == (c.n * (c.n + 1) // 2,)
    W = synth.standardized(d["bed_rows"][d["keep"]], c.n, d["freq"])
    G = (W @ W.T) / W.shape[1]
    ref = G[np.tril_indices(c.n)]
    assert np.max(np.abs(lt - ref)) <= 2.0 * np.finfo(np.float32).eps * np.max(np.abs(G))
    # the written GRM drives the --grm path to the --geno solution
    mm = ["--minMAF", str(float(d["min_maf"])), "--maxMissingRate", str(float(d["max_miss"]))]
    res = {}
    for kind, src in (("geno", ["--geno", p["geno"], *mm]), ("grm", ["--grm", out])):
        o = os.path.join(p["dir"], "rt_" + kind)
        r = run([PCG, "--input", p["input"], "--output", o, *SIG, *src])
        assert r.returncode == 0, r.stderr + r.stdout
        res[kind] = {s: synth.read_vector_bin(f"{o}_{s}.bin") for s in ("beta", "uhat", "bhat")}
    for s in ("beta", "uhat", "bhat"):
        assert rel(res["grm"][s], res["geno"][s]) < 1e-5, s
