"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker. The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build(force: bool = False) -> str:
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])
    return path


def lib():
    global _LIB
    if _LIB is None:
        _LIB = C.CDLL(build())
        _declare(_LIB)
    return _LIB


class OrcModel(C.Structure):
    _fields_ = [
        ("nrec", C.c_int64), ("n", C.c_int32), ("p", C.c_int32),
        ("indiv", C.c_void_p), ("times", C.c_void_p), ("X", C.c_void_p),
        ("use_grm", C.c_int32), ("geno", C.c_void_p), ("m", C.c_int64), ("pitch", C.c_int64),
        ("freq", C.c_void_p), ("grm_lt", C.c_void_p), ("theta", C.c_double * 5),
    ]


class PcgOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int32), ("x0_mode", C.c_int32),
                ("rule", C.c_int32), ("stagnation", C.c_int32)]


class PcgStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("stalled", C.c_int32), ("max_ratio", C.c_double)]


P = C.c_void_p


def _declare(L):
    L.orc_decode.argtypes = [C.c_uint8]
    L.orc_decode.restype = C.c_int
    L.orc_bed_counts.argtypes = [P, C.c_int64, C.c_int32, C.c_int64, P, P]
    L.orc_bed_qc.argtypes = [P, C.c_int64, C.c_int32, C.c_int64, C.c_float, C.c_float, P, P, P]
    L.orc_bed_qc.restype = C.c_int64
    L.orc_inv_std.argtypes = [C.c_float]
    L.orc_inv_std.restype = C.c_double
    L.orc_genetic_diag.argtypes = [P, P]
    L.orc_vx.argtypes = [P, P, C.c_int, P, C.c_int]
    L.orc_vx_dense.argtypes = [P, P, C.c_int, P]
    L.orc_geno_Wa.argtypes = [P, P, P]
    L.orc_vx_ref32.argtypes = [P, P, C.c_int, P, C.c_int]
    L.orc_vx_vc_ref32.argtypes = [P, P, C.c_int, P]
    L.orc_blocks.argtypes = [P, P]
    L.orc_blocks.restype = C.c_int64
    L.orc_precond_size.argtypes = [P]
    L.orc_precond_size.restype = C.c_int64
    L.orc_precond_build.argtypes = [P, P, P, C.c_int]
    L.orc_precond_apply.argtypes = [P, P, P, C.c_int, P]
    L.orc_pcg.argtypes = [P, P, P, C.c_int, P, P, P, P, P]
    L.orc_inverse.argtypes = [C.c_int, P, P]
    L.orc_inverse.restype = C.c_int
    L.orc_gls_blup.argtypes = [P, P, P, P, P, P, P, P, P, P, P]
    L.orc_bhat.argtypes = [P, P, P]
    L.orc_f_theta.argtypes = [P, P, P, P]
    L.orc_tildeV_solve.argtypes = [P, P, C.c_int, P]
    L.orc_sample_tildeV.argtypes = [P, P, P, P, P]
    L.orc_aireml_eval.argtypes = [P, P, P, P, P, C.c_int, P, C.c_int, P, P, P, P]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _cm(a):
    """[rows, k] -> column-major contiguous buffer (Fortran order)."""
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


class Oracle:
    """FP64 oracle model over one cohort + genetics (genotypes or GRM)."""

    def __init__(self, indiv, times, X, theta, *, geno_rows=None, freq=None, n=None,
                 grm_lt=None):
        self.indiv = np.ascontiguousarray(indiv, dtype=np.int32)
        self.times = np.ascontiguousarray(times, dtype=np.float64)
        self.X = _cm(X)
        self.nrec = int(self.indiv.shape[0])
        self.n = int(n if n is not None else self.indiv.max() + 1)
        self.p = int(self.X.shape[1] - 1)
        m = OrcModel()
        m.nrec, m.n, m.p = self.nrec, self.n, self.p
        m.indiv, m.times, m.X = _ptr(self.indiv), _ptr(self.times), _ptr(self.X)
        if grm_lt is not None:
            self.grm_lt = np.ascontiguousarray(grm_lt, dtype=np.float32)
            m.use_grm = 1
            m.grm_lt = _ptr(self.grm_lt)
        else:
            self.geno = np.ascontiguousarray(geno_rows, dtype=np.uint8)
            self.freq = np.ascontiguousarray(freq, dtype=np.float32)
            m.use_grm = 0
            m.geno = _ptr(self.geno)
            m.m = int(self.geno.shape[0])
            m.pitch = int(self.geno.shape[1]) if self.geno.ndim == 2 else (self.n + 3) // 4
            m.freq = _ptr(self.freq)
        self._m = m
        self.set_theta(theta)

    @property
    def ref(self):
        return C.byref(self._m)

    def set_theta(self, theta):
        for i in range(5):
            self._m.theta[i] = float(theta[i])
        self.theta = tuple(float(t) for t in theta)

    def genetic_diag(self):
        out = np.zeros(self.n)
        lib().orc_genetic_diag(self.ref, _ptr(out))
        return out

    def vx(self, x, grm_only=False):
        x = _cm(x if np.ndim(x) == 2 else np.asarray(x)[:, None])
        k = x.shape[1]
        y = np.zeros((self.n if grm_only else self.nrec, k), order="F")
        lib().orc_vx(self.ref, _ptr(x), k, _ptr(y), int(grm_only))
        return y

    def vx_dense(self, x):
        x = _cm(x if np.ndim(x) == 2 else np.asarray(x)[:, None])
        k = x.shape[1]
        y = np.zeros((self.nrec, k), order="F")
        lib().orc_vx_dense(self.ref, _ptr(x), k, _ptr(y))
        return y

    def vx_ref32(self, x, grm_only=False):
        x = np.asfortranarray(x if np.ndim(x) == 2 else np.asarray(x)[:, None], dtype=np.float32)
        k = x.shape[1]
        y = np.zeros((self.n if grm_only else self.nrec, k), order="F", dtype=np.float32)
        lib().orc_vx_ref32(self.ref, _ptr(x), k, _ptr(y), int(grm_only))
        return y

    def vx_vc_ref32(self, x):
        """FP32 record-pair restatement of computeVx_multi_VC (varCompEst.cpp:322-352):
        the CPU timing baseline of runAIREML's operator (GRM models only)."""
        x = np.asfortranarray(np.asarray(x, dtype=np.float32).reshape(self.nrec, -1))
        k = x.shape[1]
        y = np.zeros((self.nrec, k), dtype=np.float32, order="F")
        lib().orc_vx_vc_ref32(self.ref, _ptr(x), k, _ptr(y))
        return y

    def geno_Wa(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        out = np.zeros(self.n)
        lib().orc_geno_Wa(self.ref, _ptr(a), _ptr(out))
        return out

    def precond(self, no_grm=False):
        gd = self.genetic_diag()
        q = np.zeros(max(1, lib().orc_precond_size(self.ref)))
        lib().orc_precond_build(self.ref, _ptr(gd), _ptr(q), int(no_grm))
        return q

    def precond_apply(self, q, r):
        r = _cm(r if np.ndim(r) == 2 else np.asarray(r)[:, None])
        z = np.zeros_like(r, order="F")
        lib().orc_precond_apply(self.ref, _ptr(q), _ptr(r), r.shape[1], _ptr(z))
        return z

    def pcg(self, b, *, tol=1e-5, max_iter=1000, x0_mode=0, rule=0, stagnation=True,
            x0=None, q=None):
        b = _cm(b if np.ndim(b) == 2 else np.asarray(b)[:, None])
        k = b.shape[1]
        q = self.precond() if q is None else q
        x = np.zeros_like(b, order="F") if x0 is None else _cm(x0).copy(order="F")
        o = PcgOpts(tol, max_iter, x0_mode, rule, int(stagnation))
        st = PcgStats()
        ratio = np.zeros(k)
        its = np.zeros(k, dtype=np.int32)
        lib().orc_pcg(self.ref, _ptr(q), _ptr(b), k, _ptr(x), C.byref(o), C.byref(st),
                      _ptr(ratio), _ptr(its))
        return x, {"iters": st.iters, "stalled": st.stalled, "max_ratio": st.max_ratio,
                   "col_ratio": ratio, "col_iters": its}

    def gls_blup(self, y, *, tol=1e-5, max_iter=1000, x0_mode=0, rule=0, stagnation=True):
        q = self.precond()
        y = np.ascontiguousarray(y, dtype=np.float64)
        beta = np.zeros(self.p + 1)
        vr = np.zeros(self.nrec)
        uh = np.zeros(self.n)
        bh = np.zeros(2 * self.n)
        vx = np.zeros((self.nrec, self.p + 1), order="F")
        xi = np.zeros((self.p + 1, self.p + 1))
        o = PcgOpts(tol, max_iter, x0_mode, rule, int(stagnation))
        st = PcgStats()
        lib().orc_gls_blup(self.ref, _ptr(q), _ptr(y), C.byref(o), _ptr(beta), _ptr(vr),
                           _ptr(uh), _ptr(bh), _ptr(vx), _ptr(xi), C.byref(st))
        return {"beta": beta, "vinv_r": vr, "uhat": uh, "bhat": bh, "vinv_x": vx,
                "xtvx_inv": xi, "iters": st.iters, "stalled": st.stalled}

    def f_theta(self, py, uhat):
        f = np.zeros((self.nrec, 5), order="F")
        lib().orc_f_theta(self.ref, _ptr(np.ascontiguousarray(py, dtype=np.float64)),
                          _ptr(np.ascontiguousarray(uhat, dtype=np.float64)), _ptr(f))
        return f

    def tildeV_solve(self, yv):
        yv = _cm(yv if np.ndim(yv) == 2 else np.asarray(yv)[:, None])
        out = np.zeros_like(yv, order="F")
        lib().orc_tildeV_solve(self.ref, _ptr(yv), yv.shape[1], _ptr(out))
        return out

    def sample_tildeV(self, z1, z2, z3):
        out = np.zeros(self.nrec)
        lib().orc_sample_tildeV(self.ref, _ptr(np.ascontiguousarray(z1, dtype=np.float64)),
                                _ptr(np.ascontiguousarray(z2, dtype=np.float64)),
                                _ptr(np.ascontiguousarray(z3, dtype=np.float64)), _ptr(out))
        return out

    def aireml_eval(self, y, probes, weights, probe_x0=None, *, tol=1e-5, max_iter=1000, nvc=5):
        probes = _cm(probes)
        weights = _cm(weights)
        nmc = probes.shape[1]
        x0 = None if probe_x0 is None else _cm(probe_x0)
        grad = np.zeros(5)
        ai = np.zeros(nvc * nvc)
        first = np.zeros(5)
        tr = np.zeros(5)
        o = PcgOpts(tol, max_iter, 1, 1, 0)
        lib().orc_aireml_eval(self.ref, _ptr(np.ascontiguousarray(y, dtype=np.float64)),
                              _ptr(probes), _ptr(weights), _ptr(x0), nmc, C.byref(o), nvc,
                              _ptr(grad), _ptr(ai), _ptr(first), _ptr(tr))
        return {"grad": grad[:nvc], "ai": ai.reshape(nvc, nvc), "first": first[:nvc],
                "trace": tr[:nvc]}

    def aireml_eval_reduced(self, y, probes, weights, probe_x0=None, *, tol=1e-5, max_iter=1000):
        """The reduced model's computeDerivative_first + estimateTrace_PdV +
        computeDerivative_AImat (varCompEst_reducedModel_noTime.cpp:592-705): V = sig2g A
        + sig2e I (computeV_ij :288-296) is this oracle's V on a cohort of one record per
        individual with sig2b0 = sigb01 = sig2b1 = 0; its f_theta columns for sig2e and
        sig2g are the full model's first two (V^-1 r and A V^-1 r, :524-548), so the
        gradient, trace and AI terms are orc_aireml_eval's over those two columns."""
        if self.nrec != self.n or any(self.theta[j] != 0.0 for j in (2, 3, 4)):
            raise ValueError("reduced model: one record per individual and sig2b* = 0")
        return self.aireml_eval(y, probes, weights, probe_x0, tol=tol, max_iter=max_iter, nvc=2)


def aireml_fit_reduced(orc, y, theta0, *, max_iter=50, tol=1e-3, pcg_tol=1e-10):
    """The reduced model's AI-REML loop with exact trace terms: main (:967-1099) and
    oneIter_AIREML (:784-964) of varCompEst_reducedModel_noTime.cpp, with the unit probes
    sqrt(N) e_s (s < N) as the Monte Carlo sample, so every gradient is exact. The step is
    the Newton step AI^-1 g when it is feasible (the usual case: Delta starts at 1e20,
    :1024) and otherwise the minimiser of the same convex quadratic model under the
    reference's box, signed trust-region and positivity constraints (SLSQP). Restarts
    after > 5 rejections draw rand() in the reference and are not replayed (an error).
    Returns theta, the accepted theta after every iteration, and the iteration count."""
    from scipy.optimize import minimize

    n = orc.nrec
    probes = np.sqrt(n) * np.eye(n)

    def evaluate(th):
        orc.set_theta((th[0], th[1], 0.0, 0.0, 0.0))
        e = orc.aireml_eval_reduced(y, probes, probes, tol=pcg_tol)
        return e["grad"], e["ai"]

    def step(g, A, th, delta):
        lb = np.maximum(-10.0, 1e-5 - th)
        try:
            p = np.linalg.solve(A, g)
            if (np.all(np.diag(A) > 0) and np.all(p >= lb) and np.all(p <= 10.0)
                    and np.diag(A) @ p <= delta):
                return p
        except np.linalg.LinAlgError:
            pass
        cons = [{"type": "ineq", "fun": lambda q: delta - np.diag(A) @ q}]
        r = minimize(lambda q: -(g @ q - 0.5 * q @ A @ q), np.zeros(2), method="SLSQP",
                     bounds=list(zip(lb, [10.0, 10.0])), constraints=cons,
                     options={"ftol": 1e-15, "maxiter": 1000})
        return r.x

    th = np.array(theta0, dtype=np.float64)
    delta_tr, dangerous, num_rej, gnt = 1e20, False, 0, 0.0
    traj = []
    g = A = None
    it_done = 0
    for it in range(max_iter):
        if it == 0:
            g, A = evaluate(th)
        if it == 1:
            gnt = 0.1 * np.linalg.norm(g)
        if num_rej > 5:
            raise RuntimeError("restart (rand()) not replayed")
        th_old = th.copy()
        p = step(g, A, th, delta_tr)
        scale, accepted, converged, red = 1.0, False, False, 0
        while not accepted and red <= 10:
            sp = scale * p
            tn = th + sp
            if np.any(tn < 1e-6) or np.any(tn > 10.0):
                scale *= 0.5
                red += 1
                continue
            gn, An = evaluate(tn)
            dll = sp @ (0.5 * (g + gn))
            pred = g @ sp - 0.5 * sp @ A @ sp
            if scale == 1.0 and pred < 1e-2:
                dangerous, converged = False, True
                break
            if np.linalg.norm(g) > gnt and np.linalg.norm(gn) > 2 * np.linalg.norm(g):
                scale *= 0.5
                red += 1
                continue
            rho = dll / pred
            npd = float(np.sum(np.abs(np.diag(A) * sp)))
            if rho < 0.25:
                delta_tr = 0.25 * npd
            elif rho > 0.99:
                delta_tr = max(delta_tr, 3.5 * npd)
            if rho > 0.25:
                th, g, A = tn, gn, An
                num_rej, accepted = 0, True
            else:
                scale *= 0.5
                red += 1
        if not accepted and not converged:
            delta_tr *= 0.75
            dangerous = True
            num_rej += 1
        traj.append(th.copy())
        it_done = it + 1
        if np.linalg.norm(th - th_old) / np.linalg.norm(th_old) < tol and not dangerous:
            break
        dangerous = False
        if delta_tr < 1e-4:
            break
    return th, np.array(traj), it_done


def oracle_from_synth(d, theta=None) -> Oracle:
    c = d.cohort
    theta = d.theta if theta is None else theta
    if d.grm_lt is not None:
        return Oracle(c.indiv, c.times, c.X, theta, grm_lt=d.grm_lt, n=c.n)
    rows = d.bed_rows[d.keep]
    return Oracle(c.indiv, c.times, c.X, theta, geno_rows=rows, freq=d.freq, n=c.n)


def bed_counts(rows, n):
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    m = rows.shape[0]
    alt = np.zeros(m, dtype=np.int64)
    miss = np.zeros(m, dtype=np.int64)
    lib().orc_bed_counts(_ptr(rows), m, n, rows.shape[1], _ptr(alt), _ptr(miss))
    return alt, miss


def bed_qc(rows, n, min_maf, max_miss):
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    m = rows.shape[0]
    keep = np.zeros(m, dtype=np.uint8)
    maf = np.zeros(m, dtype=np.float32)
    mr = np.zeros(m, dtype=np.float32)
    kept = lib().orc_bed_qc(_ptr(rows), m, n, rows.shape[1], min_maf, max_miss,
                            _ptr(keep), _ptr(maf), _ptr(mr))
    return keep.astype(bool), maf, mr, int(kept)


# ------------------------------------------------------------------ probe generator
def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC'11) on numpy arrays of counter
    words (broadcast), key (k0, k1): the generator of the device probes. Pinned by the
    published known-answer vectors (tests/test_oracle.py)."""
    m = np.uint64(0xFFFFFFFF)
    c = [np.asarray(x, dtype=np.uint64) & m for x in (c0, c1, c2, c3)]
    k0, k1 = np.uint64(k0) & m, np.uint64(k1) & m
    for _ in range(10):
        p0 = np.uint64(0xD2511F53) * c[0]
        p1 = np.uint64(0xCD9E8D57) * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ k0) & m, p1 & m, ((p0 >> np.uint64(32)) ^ c[3] ^ k1) & m, p0 & m]
        k0 = (k0 + np.uint64(0x9E3779B9)) & m
        k1 = (k1 + np.uint64(0xBB67AE85)) & m
    return c


def probes(kind, indiv, times, theta, nmc, seed):
    """The AI-REML probes as the device draws them, restated: kind 0 Gaussian tilde-V
    (sample_Yvs, varCompEst.cpp:775-797, with Box-Muller normals from the Philox stream),
    1 Rademacher, 2 unit vectors sqrt(nrec) e_s. Returns nrec x nmc."""
    indiv = np.asarray(indiv, dtype=np.int64)
    times = np.asarray(times, dtype=np.float64)
    nrec = indiv.shape[0]
    k0, k1 = seed & 0xFFFFFFFF, seed >> 32
    r = np.arange(nrec, dtype=np.uint64)
    out = np.zeros((nrec, nmc), order="F")
    se, sg, sb0, sb01, sb1 = theta
    l00 = np.sqrt(sb0) if sb0 > 0 else 0.0
    l10 = sb01 / l00 if sb0 > 0 else 0.0
    d11 = max(sb1 - (sb01 * sb01 / sb0 if sb0 > 0 else 0.0), 0.0)
    l11 = np.sqrt(d11)
    unit = lambda w: (w.astype(np.float64) + 0.5) * 2.0 ** -32  # noqa: E731
    for s in range(nmc):
        if kind == 1:
            w = philox4x32_10(r, r >> np.uint64(32), s, 0, k0, k1)
            out[:, s] = np.where(w[0] >> np.uint64(31), 1.0, -1.0)
        elif kind == 2:
            out[s, s] = np.sqrt(nrec)
        else:
            iu = indiv.astype(np.uint64)
            a = philox4x32_10(iu, iu >> np.uint64(32), s, 1, k0, k1)
            b = philox4x32_10(r, r >> np.uint64(32), s, 2, k0, k1)
            ra = np.sqrt(-2.0 * np.log(unit(a[0])))
            z1 = ra * np.cos(2 * np.pi * unit(a[1]))
            z2a = ra * np.sin(2 * np.pi * unit(a[1]))
            z2b = np.sqrt(-2.0 * np.log(unit(a[2]))) * np.cos(2 * np.pi * unit(a[3]))
            z3 = np.sqrt(-2.0 * np.log(unit(b[0]))) * np.cos(2 * np.pi * unit(b[1]))
            out[:, s] = np.sqrt(sg) * z1 + np.sqrt(se) * z3 + z2a * (l00 + l10 * times) + z2b * l11 * times
    return out


# ------------------------------------------------------- mixed-precision GRM product
def digits4(v):
    """Four balanced base-256 digits of int64 v (|v| <= 2^30), most significant first:
    v = sum_a d[a] 256^(3 - a), d[1..3] in [-128, 127] (k_grm_i8.hip, digits4)."""
    v = np.asarray(v, dtype=np.int64).copy()
    d = [None] * 4
    for a in (3, 2, 1):
        lo = ((v & 255) ^ 128) - 128
        d[a] = lo
        v = (v - lo) >> 8
    d[0] = v
    return np.stack(d)


def _exp_of(m):
    """frexp exponent e (m = f 2^e, f in [0.5, 1)), 0 where m == 0."""
    return np.where(m > 0, np.frexp(m)[1], 0).astype(np.int64)


def grm_mixed_product(lt, n, S0, digits, sig2g=1.0):
    """sig2g G S0 as the mixed-precision device product computes it, restated with exact
    integer arithmetic (k_grm_i8.hip): per-row block fixed point of the off-diagonal FP32
    entries (scale 2^e_i, e_i the frexp exponent of the row's largest |G_ij|, j != i),
    per-column block fixed point of S0, four balanced int8 digits of each 30-bit value,
    the digit pairs a, b < L with a + b <= L summed exactly by weight class over chunks of
    tile columns,
    then combined in FP64 in the device's order, with G_ii S0_i added in FP64.
    S0: n x k. Returns n x k."""
    S0 = np.asarray(S0, dtype=np.float64).reshape(n, -1)
    L = int(digits)
    G32 = np.zeros((n, n), dtype=np.float32)
    il = np.tril_indices(n)
    G32[il] = np.asarray(lt, dtype=np.float32)
    G32 = G32 + np.tril(G32, -1).T
    gdiag = np.diag(G32).astype(np.float64)
    off = G32.copy()
    np.fill_diagonal(off, np.float32(0))
    ei = _exp_of(np.abs(off).max(axis=1))
    dG = digits4(np.rint(np.ldexp(off.astype(np.float64), (30 - ei)[:, None])).astype(np.int64))
    ec = _exp_of(np.abs(S0).max(axis=0))
    dS = digits4(np.rint(np.ldexp(S0, (30 - ec)[None, :])).astype(np.int64))
    ntiles = (n + 63) // 64
    H = max(4, (ntiles + 255) // 256)
    total = np.zeros_like(S0)
    for h in range(H):
        j0, j1 = 64 * (h * ntiles // H), min(n, 64 * ((h + 1) * ntiles // H))
        part = np.zeros_like(S0)
        for w in range(L + 1):
            acc = np.zeros(S0.shape, dtype=np.int64)
            for a in range(w + 1):
                b = w - a
                if a < L and b < L and j1 > j0:
                    acc += dG[a][:, j0:j1] @ dS[b][j0:j1, :]
            part = part + np.ldexp(acc.astype(np.float64), ei[:, None] + ec[None, :] - 12 - 8 * w)
        total = total + part
    total = total + gdiag[:, None] * S0
    return sig2g * total


def aireml_fit_replay(orc, y, theta0, probes_of, step_fn, *, max_iter=50, tol=1e-3, pcg_tol=1e-10):
    """The full-model AI-REML loop (main :1347-1491 and oneIter_AIREML :1148-1344 of
    varCompEst.cpp), replayed on the oracle's FP64 evaluations (orc_aireml_eval):
    probes_of(e, theta) gives (probes, weights, PCG start or None) of the e-th gradient
    evaluation, at theta (the device fit numbers its evaluations the same way),
    step_fn(grad, ai, theta, delta) the
    constrained step. Control flow as the reference: step halving on box / covariance
    violations, the gradient-growth check, the rho test and trust-region update, the
    AI matrix at an accepted point, the relative-change stop and the radius stop.
    Restarts after > 5 rejections draw random points and are not replayed (an error).
    Returns theta, the theta after every iteration, and the iteration count."""
    th = np.array(theta0, dtype=np.float64)
    nv = 5
    eval_id = [0]

    def evaluate(t):
        orc.set_theta(tuple(t))
        Y, W, X0 = probes_of(eval_id[0], tuple(t))
        eval_id[0] += 1
        e = orc.aireml_eval(y, Y, W, X0, tol=pcg_tol)
        return e["grad"], e["ai"]

    delta_tr, dangerous, num_rej, gnt = 1e20, False, 0, 0.0
    traj, it_done = [], 0
    g = A = None
    for it in range(max_iter):
        if it == 0:
            g, A = evaluate(th)
        if it == 1:
            gnt = 0.1 * np.linalg.norm(g)
        if num_rej > 5:
            raise RuntimeError("restart not replayed")
        th_old = th.copy()
        p = step_fn(g, A, th, delta_tr)
        if p is None:
            delta_tr *= 0.75
            dangerous = True
            num_rej += 1
        else:
            scale, accepted, converged, red = 1.0, False, False, 0
            while not accepted and red <= 10:
                sp = scale * p
                tn = th + sp
                violated = False
                for j in range(nv):
                    if j != 3 and tn[j] < 1e-6:
                        violated = True
                    if tn[j] > 10.0:
                        violated = True
                if tn[3] > np.sqrt(max(tn[2], 0.0) * max(tn[4], 0.0)):
                    violated = True
                if violated:
                    scale *= 0.5
                    red += 1
                    continue
                gn, _ = evaluate(tn)
                dll = sp @ (0.5 * (g + gn))
                pred = g @ sp - 0.5 * sp @ A @ sp
                if scale == 1.0 and pred < 1e-2:
                    dangerous, converged = False, True
                    break
                if np.linalg.norm(g) > gnt and np.linalg.norm(gn) > 2 * np.linalg.norm(g):
                    scale *= 0.5
                    red += 1
                    continue
                rho = dll / pred
                npd = float(np.sum(np.abs(np.diag(A) * sp)))
                if rho < 0.25:
                    delta_tr = 0.25 * npd
                elif rho > 0.99:
                    delta_tr = max(delta_tr, 3.5 * npd)
                if rho > 0.25:
                    th, g = tn, gn
                    # AI at the accepted point (computeDerivative_AImat, :1321-1330)
                    orc.set_theta(tuple(th))
                    A = orc.aireml_eval(y, np.zeros((len(y), 0)), np.zeros((len(y), 0)), tol=pcg_tol)["ai"]
                    num_rej, accepted = 0, True
                else:
                    scale *= 0.5
                    red += 1
            if not accepted and not converged:
                delta_tr *= 0.75
                dangerous = True
                num_rej += 1
        traj.append(th.copy())
        it_done = it + 1
        if np.linalg.norm(th - th_old) / np.linalg.norm(th_old) < tol and not dangerous:
            break
        dangerous = False
        if delta_tr < 1e-4:
            break
    return th, np.array(traj), it_done


# ------------------------------------------------ REML log-likelihood (log|V| by SLQ)
# The reference has no log-likelihood (its AI-REML loop tracks predicted and
# approximate changes only, varCompEst.cpp:1221, 1277-1278). These restate the build's
# estimator, globe_reml_loglik: log|V| = log|Q| + log|Q^-1/2 V Q^-1/2|, the first exact
# over the per-individual preconditioner blocks Q_i (computePreconditioner_Block,
# PCGsolver.cpp:991-1058), the second by stochastic Lanczos quadrature (Ubaru, Chen and
# Saad, SIAM J. Matrix Anal. Appl. 38 (2017)) on the coefficients of preconditioned CG
# solves V x = b with probes b = F u, F F^T = Q.
def precond_blocks(indiv, times, gdiag, theta):
    """Dense per-individual blocks Q_i = sig2g G_ii 11^T + U_i B U_i^T + sig2e I,
    U_i = [1, t], B = [[sig2b0, sigb01], [sigb01, sig2b1]]: {individual: (records, Q_i)}."""
    se, sg, sb0, sb01, sb1 = theta
    indiv = np.asarray(indiv)
    out = {}
    for i in np.unique(indiv):
        r = np.nonzero(indiv == i)[0]
        t = np.asarray(times, dtype=np.float64)[r]
        Q = (sg * gdiag[i] + sb0) + sb01 * (t[:, None] + t[None, :]) + sb1 * np.outer(t, t)
        out[int(i)] = (r, Q + se * np.eye(r.size))
    return out


def logdet_precond(indiv, times, gdiag, theta):
    """log|Q| = sum over individuals of log|Q_i| (numpy slogdet of each dense block)."""
    tot = 0.0
    for _, Q in precond_blocks(indiv, times, gdiag, theta).values():
        s, ld = np.linalg.slogdet(Q)
        assert s > 0
        tot += ld
    return tot


def logdet_probes(kind, indiv, times, gdiag, theta, first, count, seed):
    """The log-determinant probes b = F u as the device draws them (k_probe.hip,
    k_logdet_probes), restated: F = [sqrt(sig2e) I | U_i K_i per individual], K_i the
    lower Cholesky factor of C_i = [[sig2g G_ii + sig2b0, sigb01], [sigb01, sig2b1]]
    (a negative pivot clipped to 0), so F F^T = Q. Probe s (global index first + s):
    kind 1 Rademacher u (record signs from Philox counter (r, s, 3), the two individual
    signs from bit 31 of words 0 and 1 of (i, s, 4)); kind 0 Gaussian u (Box-Muller on
    the same counters); kind 2 the columns of F (s < nrec: sqrt(sig2e) e_s, then
    U_i K_i e_a for s = nrec + 2 i + a). Returns nrec x count."""
    indiv = np.asarray(indiv, dtype=np.int64)
    t = np.asarray(times, dtype=np.float64)
    nrec = indiv.shape[0]
    se, sg, sb0, sb01, sb1 = theta
    c00 = sg * np.asarray(gdiag, dtype=np.float64)[indiv] + sb0
    k00 = np.sqrt(np.maximum(c00, 0.0))
    k10 = np.where(k00 > 0, sb01 / np.where(k00 > 0, k00, 1.0), 0.0)
    k11 = np.sqrt(np.maximum(sb1 - k10 * k10, 0.0))
    k0, k1 = seed & 0xFFFFFFFF, seed >> 32
    r = np.arange(nrec, dtype=np.uint64)
    iu = indiv.astype(np.uint64)
    unit = lambda w: (w.astype(np.float64) + 0.5) * 2.0 ** -32  # noqa: E731
    out = np.zeros((nrec, count), order="F")
    for j in range(count):
        s = first + j
        if kind == 2:
            if s < nrec:
                out[s, j] = np.sqrt(se)
            else:
                i, a = divmod(s - nrec, 2)
                sel = indiv == i
                out[sel, j] = (k00 + k10 * t)[sel] if a == 0 else (k11 * t)[sel]
            continue
        w = philox4x32_10(r, r >> np.uint64(32), s, 3, k0, k1)
        v = philox4x32_10(iu, iu >> np.uint64(32), s, 4, k0, k1)
        if kind == 1:
            u3 = np.where(w[0] >> np.uint64(31), 1.0, -1.0)
            ua = np.where(v[0] >> np.uint64(31), 1.0, -1.0)
            ub = np.where(v[1] >> np.uint64(31), 1.0, -1.0)
        else:
            u3 = np.sqrt(-2.0 * np.log(unit(w[0]))) * np.cos(2 * np.pi * unit(w[1]))
            ra = np.sqrt(-2.0 * np.log(unit(v[0])))
            ua = ra * np.cos(2 * np.pi * unit(v[1]))
            ub = ra * np.sin(2 * np.pi * unit(v[1]))
        out[:, j] = np.sqrt(se) * u3 + ua * (k00 + k10 * t) + ub * k11 * t
    return out


def slq_exact_terms(V, Q, B, control=False):
    """Per probe b (columns of B) the quadratic form the Lanczos quadrature converges
    to: u^T log(A) u with u = Q^-1/2 b and A = Q^-1/2 V Q^-1/2 (dense eigendecompositions;
    any square root of Q gives the same value, as u u^T -> Q^-1/2 b b^T Q^-1/2).
    control=True: u^T (log(A) - A + I) u, the device's estimator (tr(A - I) = 0 when Q's
    blocks are V's diagonal blocks)."""
    w, E = np.linalg.eigh(Q)
    Qh = (E / np.sqrt(w)) @ E.T
    lam, Z = np.linalg.eigh(Qh @ V @ Qh)
    U = Z.T @ (Qh @ B)
    f = np.log(lam) - (lam - 1.0) if control else np.log(lam)
    return (f[:, None] * U * U).sum(axis=0)


def pcg_lanczos(V, Qinv, b, tol, max_iter=1000):
    """Preconditioned CG on V x = b from x = 0 (numpy, FP64) returning r0.z0 and the
    coefficients alpha_0..alpha_{m-1}, beta_0..beta_{m-2} of its m iterations."""
    x = np.zeros_like(b)
    r = b.copy()
    z = Qinv @ r
    p = z.copy()
    rz = r @ z
    rho0, al, be = rz, [], []
    nb = np.linalg.norm(b)
    for _ in range(max_iter):
        vp = V @ p
        a = rz / (p @ vp)
        al.append(a)
        x += a * p
        r -= a * vp
        if np.linalg.norm(r) < tol * nb:
            break
        z = Qinv @ r
        rz2 = r @ z
        be.append(rz2 / rz)
        p = z + be[-1] * p
        rz = rz2
    return rho0, np.array(al), np.array(be)


def lanczos_t00(alpha):
    """T_00 = 1/alpha_0 = z0^T V z0 / r0.z0: the Rayleigh quotient of A at u."""
    return 1.0 / float(np.asarray(alpha)[0])


def lanczos_log_quadrature(alpha, beta):
    """e1^T log(T) e1 of the Lanczos tridiagonal of a CG solve (Saad, Iterative Methods
    for Sparse Linear Systems, 2nd ed., sec. 6.7.3): T_00 = 1/alpha_0, T_jj = 1/alpha_j +
    beta_{j-1}/alpha_{j-1}, T_{j-1,j} = sqrt(beta_{j-1})/alpha_{j-1}; Gauss quadrature
    sum_k tau_k^2 log theta_k over T's eigenpairs (scipy eigh_tridiagonal)."""
    from scipy.linalg import eigh_tridiagonal
    alpha = np.asarray(alpha, dtype=np.float64)
    beta = np.asarray(beta, dtype=np.float64)
    m = alpha.size
    d = 1.0 / alpha
    d[1:] += beta[: m - 1] / alpha[: m - 1]
    e = np.sqrt(beta[: m - 1]) / alpha[: m - 1]
    th, Z = eigh_tridiagonal(d, e)
    return float((Z[0] ** 2 * np.log(th)).sum())


def reml_loglik_dense(V, X, y):
    """The REML log-likelihood -1/2 (log|V| + log|X^T V^-1 X| + y^T P y + (n - q) log 2pi),
    P = V^-1 - V^-1 X (X^T V^-1 X)^-1 X^T V^-1, from dense V (numpy slogdet).
    Returns (loglik, log|V|, log|X^T V^-1 X|, y^T P y)."""
    Vi = np.linalg.inv(V)
    XtViX = X.T @ Vi @ X
    P = Vi - Vi @ X @ np.linalg.solve(XtViX, X.T @ Vi)
    ld1 = np.linalg.slogdet(V)[1]
    ld2 = np.linalg.slogdet(XtViX)[1]
    ypy = float(y @ P @ y)
    n, q = X.shape
    return -0.5 * (ld1 + ld2 + ypy + (n - q) * np.log(2 * np.pi)), ld1, ld2, ypy
