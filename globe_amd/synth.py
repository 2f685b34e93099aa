"""Synthetic LGMM cohorts and the reference's file formats.

Recipe (seeded numpy ``default_rng``):
  * allele frequency f_s ~ U(f_lo, f_hi); dosage ~ Binomial(2, f_s) counting the A1
    allele; each genotype missing with probability ``miss``; PLINK 1 SNP-major codes
    2 -> 00, 1 -> 10, 0 -> 11, missing -> 01 (PCGsolver.cpp:83-94).
  * records per individual 1 + Poisson(lam), contiguous, IDs 1..N; times U(0, 10);
    covariates intercept, sex ~ Bernoulli(0.5), age ~ U(40, 70); X = [covs, TIME].
  * y = X beta + u + b0 + b1 t + e, beta = (1.0, 0.3, 0.01, 0.05),
    u = sqrt(sig2g / M) W a, a ~ N(0, I); (b0, b1) ~ N(0, B); e ~ N(0, sig2e).
Every real value is rounded to FP32 because the reference parses its text input as
float (PCGsolver.cpp:559-569); the arrays returned here are exactly what a reader
of the written files sees.

File formats kept byte-compatible with the reference:
  * text input: header ``INDIV PHENO TIMES <covariates...>`` (PCGsolver.cpp:476-497)
  * ``<prefix>.bed`` magic 6C 1B 01 + SNP rows of ceil(N/4) bytes (:650-667), ``.bim``
    one line per SNP (:634-636), ``.fam`` (never read by the reference)
  * ``<prefix>.grm.bin``: N(N+1)/2 float32 lower triangle, row-major (:778-809)
  * result vectors: int32 n + n float32 (utils.cpp:5-21)
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

BETA_DEFAULT = (1.0, 0.3, 0.01, 0.05)
THETA_DEFAULT = (0.6, 0.4, 0.4, 0.2, 0.2)  # sig2e, sig2g, sig2b0, sigb01, sig2b1

# dosage (0, 1, 2, missing=-1) -> PLINK 2-bit code
_CODE_OF = {2: 0b00, 1: 0b10, 0: 0b11, -1: 0b01}
# 2-bit code -> dosage used by the operator (missing counts as 0, PCGsolver.cpp:837)
DOSAGE_OF_CODE = np.array([2, 0, 1, 0], dtype=np.int8)
ALLELE_OF_CODE = np.array([2, -1, 1, 0], dtype=np.int8)


@dataclass
class Cohort:
    n: int
    indiv: np.ndarray          # int32 [nrec], 0-based
    times: np.ndarray          # float64 [nrec] (FP32-exact)
    covars: np.ndarray         # float64 [nrec, p] (FP32-exact)
    cov_names: list = field(default_factory=lambda: ["intercept", "sex", "age"])

    @property
    def nrec(self) -> int:
        return int(self.indiv.shape[0])

    @property
    def p(self) -> int:
        return int(self.covars.shape[1])

    @property
    def X(self) -> np.ndarray:
        """[nrec, p+1] with TIME as the last column (PCGsolver.cpp:574)."""
        return np.concatenate([self.covars, self.times[:, None]], axis=1)

    def offsets(self) -> np.ndarray:
        off = np.zeros(self.n + 1, dtype=np.int64)
        np.add.at(off, self.indiv.astype(np.int64) + 1, 1)
        return np.cumsum(off)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def simulate_cohort(n: int, lam: float, rng: np.random.Generator) -> Cohort:
    counts = 1 + rng.poisson(lam, size=n)
    indiv = np.repeat(np.arange(n, dtype=np.int32), counts)
    nrec = indiv.shape[0]
    times = f32(rng.uniform(0.0, 10.0, size=nrec))
    sex = rng.binomial(1, 0.5, size=n).astype(np.float64)[indiv]
    age = f32(rng.uniform(40.0, 70.0, size=n))[indiv]
    covars = np.stack([np.ones(nrec), sex, age], axis=1)
    return Cohort(n=n, indiv=indiv, times=times, covars=covars)


def simulate_dosages(n: int, m: int, rng: np.random.Generator, f_lo=0.05, f_hi=0.5,
                     miss=0.0) -> np.ndarray:
    """int8 [m, n] dosages with -1 for missing."""
    f = rng.uniform(f_lo, f_hi, size=m)
    g = rng.binomial(2, f[:, None], size=(m, n)).astype(np.int8)
    if miss > 0:
        g[rng.random((m, n)) < miss] = -1
    return g


def pack_bed(dos: np.ndarray, phantom_code: int = 0b00) -> np.ndarray:
    """Pack [m, n] dosages into SNP-major rows of ceil(n/4) bytes. Phantom pairs of
    the last byte get ``phantom_code`` (PLINK writes 00)."""
    m, n = dos.shape
    nb = (n + 3) // 4
    codes = np.full((m, nb * 4), phantom_code, dtype=np.uint8)
    lut = np.zeros(256, dtype=np.uint8)
    for d, c in _CODE_OF.items():
        lut[d & 0xFF] = c
    codes[:, :n] = lut[dos.astype(np.int16) & 0xFF]
    codes = codes.reshape(m, nb, 4)
    return (codes[..., 0] | (codes[..., 1] << 2) | (codes[..., 2] << 4) | (codes[..., 3] << 6)).astype(np.uint8)


def unpack_codes(rows: np.ndarray, n: int) -> np.ndarray:
    """[m, nb] bytes -> [m, n] 2-bit codes (brute force, PLINK bit order)."""
    m, nb = rows.shape
    sh = np.array([0, 2, 4, 6], dtype=np.uint8)
    codes = (rows[:, :, None] >> sh[None, None, :]) & 3
    return codes.reshape(m, nb * 4)[:, :n]


def qc_reference(rows: np.ndarray, n: int, min_maf: float, max_miss: float):
    """numpy restatement of the FP32 QC decision (PCGsolver.cpp:706-743)."""
    codes = unpack_codes(rows, n)
    alt = (2 * (codes == 0) + (codes == 2)).sum(axis=1).astype(np.int64)
    miss = (codes == 1).sum(axis=1).astype(np.int64)
    n_obs = n - miss
    with np.errstate(divide="ignore", invalid="ignore"):
        maf = np.where(n_obs > 0, alt.astype(np.float32) / (np.float32(2.0) * n_obs.astype(np.float32)),
                       np.float32(0.0)).astype(np.float32)
    mr = (miss.astype(np.float32) / np.float32(n)).astype(np.float32)
    mm = np.float32(min_maf)
    keep = (maf >= mm) & (maf <= np.float32(1.0) - mm) & (mr <= np.float32(max_miss))
    return keep, maf, mr, alt, miss


def standardized(rows: np.ndarray, n: int, freq: np.ndarray) -> np.ndarray:
    """FP64 [n, m] standardized genotypes w = (g - 2f) inv_std, missing g := 0."""
    codes = unpack_codes(rows, n)
    g = DOSAGE_OF_CODE[codes].astype(np.float64)  # [m, n]
    f = freq.astype(np.float32).astype(np.float64)
    d = 2.0 * f * (1.0 - f)
    inv = np.where(d > 1e-12, 1.0 / np.sqrt(np.where(d > 0, d, 1.0)), 1.0)
    return ((g - 2.0 * f[:, None]) * inv[:, None]).T


def simulate_phenotype(c: Cohort, u: np.ndarray, theta, rng, beta=BETA_DEFAULT) -> np.ndarray:
    se, sg, sb0, sb01, sb1 = theta
    B = np.array([[sb0, sb01], [sb01, sb1]])
    L = np.linalg.cholesky(B)
    b = rng.standard_normal((c.n, 2)) @ L.T
    e = rng.standard_normal(c.nrec) * np.sqrt(se)
    y = c.X @ np.asarray(beta) + u[c.indiv] + b[c.indiv, 0] + b[c.indiv, 1] * c.times + e
    return f32(y)


@dataclass
class SynthData:
    cohort: Cohort
    y: np.ndarray
    bed_rows: np.ndarray | None = None   # all SNPs (pre-QC), [m_total, ceil(n/4)]
    keep: np.ndarray | None = None
    freq: np.ndarray | None = None       # float32 [m_kept]
    grm_lt: np.ndarray | None = None     # float32 [n(n+1)/2]
    theta: tuple = THETA_DEFAULT


def make_geno_case(n: int, m: int, lam: float, seed: int, f_lo=0.05, f_hi=0.5, miss=0.0,
                   min_maf=0.01, max_miss=0.2, theta=THETA_DEFAULT) -> SynthData:
    rng = np.random.default_rng(seed)
    c = simulate_cohort(n, lam, rng)
    dos = simulate_dosages(n, m, rng, f_lo, f_hi, miss)
    rows = pack_bed(dos)
    keep, maf, _, _, _ = qc_reference(rows, n, min_maf, max_miss)
    freq = maf[keep].astype(np.float32)
    W = standardized(rows[keep], n, freq)
    mk = int(keep.sum())
    a = rng.standard_normal(mk)
    u = np.sqrt(theta[1] / max(mk, 1)) * (W @ a)
    y = simulate_phenotype(c, u, theta, rng)
    return SynthData(cohort=c, y=y, bed_rows=rows, keep=keep, freq=freq, theta=theta)


def make_grm_case(n: int, m: int, lam: float, seed: int, f_lo=0.05, f_hi=0.5,
                  theta=THETA_DEFAULT) -> SynthData:
    rng = np.random.default_rng(seed)
    c = simulate_cohort(n, lam, rng)
    dos = simulate_dosages(n, m, rng, f_lo, f_hi, 0.0)
    rows = pack_bed(dos)
    keep, maf, _, _, _ = qc_reference(rows, n, 0.0, 1.0)
    W = standardized(rows, n, maf)
    G = (W @ W.T) / m
    il = np.tril_indices(n)
    lt = G[il].astype(np.float32)  # row-major lower triangle: i*(i+1)/2 + j
    a = rng.standard_normal(m)
    u = np.sqrt(theta[1] / m) * (W @ a)
    y = simulate_phenotype(c, u, theta, rng)
    return SynthData(cohort=c, y=y, grm_lt=lt, theta=theta)


@dataclass
class ReducedCase:
    """Reduced-model cohort (varCompEst_reducedModel_noTime.cpp): one record per
    individual in GRM order, covariates without a TIME column."""
    y: np.ndarray          # float64 [n] (FP32-exact)
    covars: np.ndarray     # float64 [n, p] (FP32-exact)
    grm_lt: np.ndarray     # float32 [n(n+1)/2]
    cov_names: list = field(default_factory=lambda: ["intercept", "sex", "age"])

    @property
    def n(self) -> int:
        return int(self.y.shape[0])


def make_reduced_case(n: int, m: int, seed: int, theta=(0.5, 0.5), f_lo=0.05, f_hi=0.5,
                      beta=(1.0, 0.3, 0.01)) -> ReducedCase:
    """y = X beta + u + e with u ~ N(0, sig2g G), e ~ N(0, sig2e I), G = W W^T / m."""
    rng = np.random.default_rng(seed)
    dos = simulate_dosages(n, m, rng, f_lo, f_hi, 0.0)
    rows = pack_bed(dos)
    _, maf, _, _, _ = qc_reference(rows, n, 0.0, 1.0)
    W = standardized(rows, n, maf)
    G = (W @ W.T) / m
    lt = G[np.tril_indices(n)].astype(np.float32)
    covars = np.stack([np.ones(n), rng.binomial(1, 0.5, size=n).astype(np.float64),
                       f32(rng.uniform(40.0, 70.0, size=n))], axis=1)
    u = np.sqrt(theta[1] / m) * (W @ rng.standard_normal(m))
    y = f32(covars @ np.asarray(beta) + u + np.sqrt(theta[0]) * rng.standard_normal(n))
    return ReducedCase(y=y, covars=covars, grm_lt=lt)


def write_reduced_input(path: str, d: ReducedCase) -> None:
    """Header INDIV PHENO <covariates> (readInputData, varCompEst_reducedModel_noTime.cpp:191-248)."""
    with open(path, "w") as fh:
        fh.write(" ".join(["INDIV", "PHENO", *d.cov_names]) + "\n")
        for i in range(d.n):
            vals = [str(i + 1), repr(float(np.float32(d.y[i])))]
            vals += [repr(float(np.float32(v))) for v in d.covars[i]]
            fh.write(" ".join(vals) + "\n")


def grm_dense(lt: np.ndarray, n: int) -> np.ndarray:
    G = np.zeros((n, n))
    il = np.tril_indices(n)
    G[il] = lt
    return G + np.tril(G, -1).T


# --------------------------------------------------------------------------- files
def write_text_input(path: str, c: Cohort, y: np.ndarray) -> None:
    cols = ["INDIV", "PHENO", "TIMES"] + list(c.cov_names)
    with open(path, "w") as fh:
        fh.write(" ".join(cols) + "\n")
        ids = c.indiv.astype(np.int64) + 1
        y32 = y.astype(np.float32)
        t32 = c.times.astype(np.float32)
        cv = c.covars.astype(np.float32)
        for r in range(c.nrec):
            vals = [str(int(ids[r])), repr(float(y32[r])), repr(float(t32[r]))]
            vals += [repr(float(v)) for v in cv[r]]
            fh.write(" ".join(vals) + "\n")


def write_plink(prefix: str, rows: np.ndarray, n: int) -> None:
    m = rows.shape[0]
    with open(prefix + ".bed", "wb") as fh:
        fh.write(bytes([0x6C, 0x1B, 0x01]))
        fh.write(np.ascontiguousarray(rows, dtype=np.uint8).tobytes())
    with open(prefix + ".bim", "w") as fh:
        for s in range(m):
            fh.write(f"1\trs{s + 1}\t0\t{s + 1}\tA\tG\n")
    with open(prefix + ".fam", "w") as fh:
        for i in range(n):
            fh.write(f"{i + 1} {i + 1} 0 0 0 -9\n")


def write_grm_bin(prefix: str, lt: np.ndarray) -> None:
    with open(prefix + ".grm.bin", "wb") as fh:
        fh.write(np.ascontiguousarray(lt, dtype=np.float32).tobytes())


def read_vector_bin(path: str) -> np.ndarray:
    with open(path, "rb") as fh:
        n = int(np.frombuffer(fh.read(4), dtype=np.int32)[0])
        v = np.frombuffer(fh.read(4 * n), dtype=np.float32)
    if v.shape[0] != n:
        raise ValueError(f"{path}: short vector file")
    return v.astype(np.float64)


def write_case_files(d: SynthData, dirpath: str, name: str = "case") -> dict:
    os.makedirs(dirpath, exist_ok=True)
    paths = {"input": os.path.join(dirpath, name + ".txt")}
    write_text_input(paths["input"], d.cohort, d.y)
    if d.bed_rows is not None:
        paths["geno"] = os.path.join(dirpath, name)
        write_plink(paths["geno"], d.bed_rows, d.cohort.n)
    if d.grm_lt is not None:
        paths["grm"] = os.path.join(dirpath, name)
        write_grm_bin(paths["grm"], d.grm_lt)
    return paths
