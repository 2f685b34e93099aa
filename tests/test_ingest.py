"""Genotype ingest (prepGenotypeInput, PCGsolver.cpp:615-763) through the pinned,
chunked pipeline with the QC counts, FP32 keep rule and compaction on the device.
From a PLINK file set or a host buffer, and whatever the chunk size (one row per
chunk, a few rows, the default), the counts, keep set and frequencies must equal the
FP32 QC restatement bit for bit, and the resident rows must be the kept rows."""
import numpy as np
import pytest

import globe_amd as G
from globe_amd import synth
from helpers import load_golden

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _write_plink(prefix, rows, fam_n=None, bim_m=None):
    with open(prefix + ".bed", "wb") as fh:
        fh.write(bytes([0x6C, 0x1B, 0x01]))
        fh.write(np.ascontiguousarray(rows).tobytes())
    with open(prefix + ".bim", "w") as fh:
        for s in range(rows.shape[0] if bim_m is None else bim_m):
            fh.write(f"1 rs{s} 0 {s + 1} A G\n")
    if fam_n is not None:
        with open(prefix + ".fam", "w") as fh:
            for i in range(fam_n):
                fh.write(f"{i + 1} {i + 1} 0 0 0 -9\n")


def _pack_codes(codes):
    """[m, 4 nb] 2-bit codes -> [m, nb] bytes (PLINK bit order)."""
    c = codes.reshape(codes.shape[0], -1, 4).astype(np.uint8)
    return (c[..., 0] | (c[..., 1] << 2) | (c[..., 2] << 4) | (c[..., 3] << 6)).astype(np.uint8)


def _context(d, ctx_factory):
    g = ctx_factory()
    c = d["cohort"]
    g.set_cohort(c.indiv, c.times, c.X)
    return g


def _expect_rows(rows, keep, n):
    # the device holds dosages, so a missing call (01) reads back as 11 (dosage 0)
    e = synth.unpack_codes(rows[keep], n)
    e[e == 1] = 3
    return e


@pytest.mark.parametrize("name", ["geno_n150_odd", "geno_qc_n257"])
@pytest.mark.parametrize("chunk", [1, 448, None])
def test_file_ingest_matches_qc_and_buffer(name, chunk, tmp_path, gpu_ctx_factory):
    _say("ingest-file", name, chunk)
    d = load_golden(name)
    n = d["cohort"].n
    rows = d["bed_rows"]
    assert rows.shape[1] == (n + 3) // 4
    mn, mx = float(d["min_maf"]), float(d["max_miss"])
    prefix = str(tmp_path / name)
    _write_plink(prefix, rows, fam_n=n)
    g = _context(d, gpu_ctx_factory)
    if chunk:
        g.set_ingest_chunk(chunk)
    kept = g.load_bed_file(prefix, mn, mx)
    st = g.ingest_stats()
    per_chunk = max(1, (chunk or (256 << 20)) // rows.shape[1])
    assert st["chunks"] == -(-rows.shape[0] // per_chunk)
    keep, maf, _, alt, miss = synth.qc_reference(rows, n, mn, mx)
    q = g.qc()
    assert np.array_equal(q["alt"], alt) and np.array_equal(q["miss"], miss)
    assert np.array_equal(q["keep"], keep) and np.array_equal(q["maf"], maf)
    assert g.m_total == rows.shape[0] and kept == int(keep.sum()) == int(d["ref_kept"])
    assert np.array_equal(g.allele_freqs(), maf[keep])
    got = g.geno_rows()
    nb = (n + 3) // 4
    assert np.array_equal(synth.unpack_codes(got[:, :nb], n), _expect_rows(rows, keep, n))
    # the host-buffer path (default chunk) stages the same resident state
    b = _context(d, gpu_ctx_factory)
    b.prepGenotypeInput(rows, mn, mx)
    assert np.array_equal(b.geno_rows(), got)
    assert np.array_equal(b.computeGeneticDiagonal(), g.computeGeneticDiagonal())


@pytest.mark.parametrize("chunk", [1, 130, 1000])
def test_buffer_ingest_chunked_ragged(chunk, gpu_ctx_factory):
    # random codes (missing included, and in the phantom fields of the last byte) over
    # a ragged N, thresholds that keep part of the rows, chunk boundaries everywhere
    _say("ingest-buffer", chunk)
    n = 301
    rng = np.random.default_rng(chunk)
    c = synth.simulate_cohort(n, 1.0, rng)
    nb = (n + 3) // 4
    codes = rng.choice(4, size=(257, 4 * nb), p=[0.3, 0.05, 0.35, 0.3]).astype(np.uint8)
    codes[::5] = rng.choice(4, size=(len(codes[::5]), 4 * nb), p=[0.01, 0.3, 0.04, 0.65])
    rows = _pack_codes(codes)
    g = gpu_ctx_factory()
    g.set_cohort(c.indiv, c.times, c.X)
    g.set_ingest_chunk(chunk)
    g.prepGenotypeInput(rows, 0.02, 0.1)
    keep, maf, _, alt, miss = synth.qc_reference(rows, n, 0.02, 0.1)
    assert 0 < keep.sum() < rows.shape[0]
    q = g.qc()
    assert np.array_equal(q["alt"], alt) and np.array_equal(q["miss"], miss)
    assert np.array_equal(q["keep"], keep) and np.array_equal(q["maf"], maf)
    got = g.geno_rows()
    assert np.array_equal(synth.unpack_codes(got[:, :nb], n), _expect_rows(rows, keep, n))


def test_file_ingest_errors(tmp_path, gpu_ctx_factory):
    _say("ingest-errors")
    d = load_golden("geno_n150_odd")
    n = d["cohort"].n
    rows = d["bed_rows"]
    g = _context(d, gpu_ctx_factory)
    with pytest.raises(G.GlobeError) as e:  # no such file set
        g.load_bed_file(str(tmp_path / "missing"))
    assert e.value.code == 6
    pre = str(tmp_path / "fam")
    _write_plink(pre, rows, fam_n=n + 1)
    with pytest.raises(G.GlobeError) as e:  # .fam disagrees with the cohort
        g.load_bed_file(pre)
    assert e.value.code == 1 and ".fam" in str(e.value)
    pre = str(tmp_path / "short")
    _write_plink(pre, rows[:-1], bim_m=rows.shape[0])
    with pytest.raises(G.GlobeError) as e:  # .bed shorter than the .bim says
        g.load_bed_file(pre)
    assert e.value.code == 1 and "fewer than" in str(e.value)
    pre = str(tmp_path / "ok")
    _write_plink(pre, rows)  # no .fam: accepted, as the reference never reads it
    assert g.load_bed_file(pre, float(d["min_maf"]), float(d["max_miss"])) == int(d["ref_kept"])
