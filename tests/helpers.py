"""Shared test helpers: golden fixtures and synthetic cases for oracle and device."""
import os

import numpy as np

from globe_amd import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    c = synth.Cohort(n=int(d["indiv"].max()) + 1, indiv=d["indiv"].astype(np.int32),
                     times=d["times"], covars=d["covars"])
    d["cohort"] = c
    if "bed_rows" in d:
        keep, maf, _, _, _ = synth.qc_reference(d["bed_rows"], c.n, float(d["min_maf"]),
                                                float(d["max_miss"]))
        d["keep"] = keep
        d["freq"] = maf[keep].astype(np.float32)
    return d


def golden_oracle(d, theta=None):
    import oracle as O
    c = d["cohort"]
    theta = tuple(d["theta"]) if theta is None else theta
    if "grm_lt" in d:
        return O.Oracle(c.indiv, c.times, c.X, theta, grm_lt=d["grm_lt"], n=c.n)
    return O.Oracle(c.indiv, c.times, c.X, theta, geno_rows=d["bed_rows"][d["keep"]],
                    freq=d["freq"], n=c.n)


def dense_V(c, G, theta):
    se, sg, sb0, sb01, sb1 = theta
    J = np.zeros((c.nrec, c.n))
    J[np.arange(c.nrec), c.indiv] = 1.0
    t = c.times
    same = c.indiv[:, None] == c.indiv[None, :]
    T = same * (sb0 + sb01 * (t[:, None] + t[None, :]) + sb1 * np.outer(t, t))
    return sg * J @ G @ J.T + T + se * np.eye(c.nrec), J
