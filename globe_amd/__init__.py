"""globe_amd: MI355X-native matrix-free LGMM solve (ctypes binding of libglobe_amd.so).

The product is the C ABI in include/globe.h (libglobe_amd.so, HIP kernels for
gfx950) and the C++ CLIs built on it (globe_amd/bin/runPCGsolver, runAIREML).
This module is the Python binding used by the tests and bench.py; its method
names follow the reference's C++ functions so parity tests read like the
reference's own call sites.

There is no CPU fallback: if libglobe_amd.so is missing or no GPU is visible the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libglobe_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "globe.h")

GLOBE_OK = 0
VX_FULL, VX_GRM_ONLY = 0, 1
PREC_FP64, PREC_MIXED, PREC_FP64_VALU = 0, 1, 2
RULE_MAX_RATIO, RULE_ALL_COLUMNS, RULE_PER_COLUMN = 0, 1, 2
PCG_FIXED_DIGITS = 1  # GLOBE_PCG_FIXED_DIGITS: mixed precision keeps 4 digits
X0_ZERO, X0_PRECOND, X0_GIVEN = 0, 1, 2

_LIB = None


class GlobeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"globe error {code}: {msg}")
        self.code = code


class PcgOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int32), ("x0_mode", C.c_int32),
                ("rule", C.c_int32), ("stagnation", C.c_int32), ("precision", C.c_int32),
                ("flags", C.c_int32)]


class PcgStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("stalled", C.c_int32), ("max_ratio", C.c_double),
                ("seconds", C.c_double)]


class KernelStats(C.Structure):
    _fields_ = [("launches", C.c_int64), ("ms_total", C.c_double), ("bytes_per_launch", C.c_double),
                ("flops_per_launch", C.c_double), ("issued_ops_per_launch", C.c_double)]


class AiremlOpts(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("tol", C.c_double), ("num_mc", C.c_int32),
                ("probes", C.c_int32), ("seed", C.c_uint64), ("pcg", PcgOpts)]


class Loglik(C.Structure):
    _fields_ = [("loglik", C.c_double), ("logdet_v", C.c_double), ("logdet_v_se", C.c_double),
                ("logdet_q", C.c_double), ("logdet_xtvx", C.c_double), ("ypy", C.c_double),
                ("probes", C.c_int32), ("max_steps", C.c_int32), ("seconds", C.c_double)]


class Request(C.Structure):
    _fields_ = [("nrec", C.c_int64), ("indiv", C.c_void_p), ("times", C.c_void_p), ("y", C.c_void_p),
                ("theta", C.c_double * 5)]


class ServeOpts(C.Structure):
    _fields_ = [("slots", C.c_int32), ("max_records", C.c_int64), ("hbm_budget", C.c_int64),
                ("tol", C.c_double), ("max_iter", C.c_int32), ("precision", C.c_int32)]


class ServeResult(C.Structure):
    _fields_ = [("id", C.c_int64), ("iters", C.c_int32), ("converged", C.c_int32), ("ratio", C.c_double),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


LOG_FN = C.CFUNCTYPE(None, C.c_char_p, C.c_void_p)


def header_symbols() -> list[str]:
    """Every function declared in include/globe.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(globe_[a-z0-9_]+)\s*\(", txt)))


def load_library(path: str = LIB_PATH):
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run __graft_entry__.build() or python globe_amd/build.py")
    L = C.CDLL(path)
    P = C.c_void_p
    i64, i32, dbl = C.c_int64, C.c_int32, C.c_double
    sig = {
        "globe_abi_version": ([], C.c_int),
        "globe_device_count": ([], C.c_int),
        "globe_create": ([C.c_int, C.POINTER(P)], C.c_int),
        "globe_destroy": ([P], None),
        "globe_last_error": ([P], C.c_char_p),
        "globe_set_cohort": ([P, i64, i32, i32, P, P, P, i64], C.c_int),
        "globe_load_bed": ([P, P, i64, i64, C.c_float, C.c_float, P], C.c_int),
        "globe_load_bed_file": ([P, C.c_char_p, C.c_float, C.c_float, P, P], C.c_int),
        "globe_set_ingest_chunk": ([P, i64], C.c_int),
        "globe_ingest_stats": ([P, P], C.c_int),
        "globe_get_qc": ([P, P, P, P, P], C.c_int),
        "globe_get_freq": ([P, P], C.c_int),
        "globe_load_grm": ([P, P], C.c_int),
        "globe_genetic_diag": ([P, P], C.c_int),
        "globe_set_varcomp": ([P, P], C.c_int),
        "globe_vx": ([P, P, i64, C.c_int, P, i64, C.c_int, C.c_int], C.c_int),
        "globe_vx_dev": ([P, P, i64, C.c_int, P, i64, C.c_int, C.c_int], C.c_int),
        "globe_vx_dev_noalloc": ([P, P, i64, C.c_int, P, i64, C.c_int, C.c_int], C.c_int),
        "globe_precond_apply": ([P, P, i64, C.c_int, P, i64], C.c_int),
        "globe_pcg": ([P, P, i64, C.c_int, P, i64, P, P, P, P], C.c_int),
        "globe_pcg_dev": ([P, P, i64, C.c_int, P, i64, P, P, P, P], C.c_int),
        "globe_gls_blup": ([P, P, P, P, P, P, P, P], C.c_int),
        "globe_aireml_eval": ([P, P, P, P, P, C.c_int, P, P, P, P, P], C.c_int),
        "globe_aireml_eval_reduced": ([P, P, P, P, P, C.c_int, P, P, P, P, P], C.c_int),
        "globe_dev_alloc": ([P, i64, C.POINTER(P)], C.c_int),
        "globe_dev_free": ([P, P], C.c_int),
        "globe_memcpy_h2d": ([P, P, P, i64], C.c_int),
        "globe_memcpy_d2h": ([P, P, P, i64], C.c_int),
        "globe_synchronize": ([P], C.c_int),
        "globe_synth_bed": ([P, i64, C.c_uint64, C.c_float, C.c_float, C.c_float, C.c_float,
                             C.c_float, P], C.c_int),
        "globe_geno_Wa": ([P, P, P], C.c_int),
        "globe_get_geno_rows": ([P, i64, i64, P, P], C.c_int),
        "globe_dev_copy": ([P, P, P, i64], C.c_int),
        "globe_profile_enable": ([P, C.c_int], C.c_int),
        "globe_force_variant": ([P, C.c_int], C.c_int),
        "globe_device_memory": ([P, P, P], C.c_int),
        "globe_variant_counts": ([P, P], C.c_int),
        "globe_set_mixed_digits": ([P, C.c_int], C.c_int),
        "globe_set_graphs": ([P, C.c_int], C.c_int),
        "globe_set_row_gather": ([P, C.c_int], C.c_int),
        "globe_profile_read": ([P, C.c_int, P], C.c_int),
        "globe_aireml_step": ([P, P, P, dbl, P], C.c_int),
        "globe_aireml_step_reduced": ([P, P, P, dbl, P], C.c_int),
        "globe_aireml_probes": ([P, C.c_int, C.c_int, C.c_uint64, P, P], C.c_int),
        "globe_probe_seed": ([C.c_uint64, C.c_uint64], C.c_uint64),
        "globe_reml_loglik": ([P, P, C.c_int, C.c_int, C.c_uint64, P, P], C.c_int),
        "globe_aireml_fit": ([P, P, P, P, P, P, P, P, P, P], C.c_int),
        "globe_aireml_fit_reduced": ([P, P, P, P, P, P, P, P, P, P], C.c_int),
        "globe_grm_from_geno": ([P], C.c_int),
        "globe_get_grm": ([P, P], C.c_int),
        "globe_get_grm_rows": ([P, i64, i64, P], C.c_int),
        "globe_serve_open": ([P, P], C.c_int),
        "globe_serve_submit": ([P, P, P], C.c_int),
        "globe_serve_step": ([P, P, P, P], C.c_int),
        "globe_serve_take": ([P, i64, P, P], C.c_int),
        "globe_serve_finished": ([P, P, i32, P], C.c_int),
        "globe_serve_close": ([P], C.c_int),
        "globe_serve_usage": ([P, P], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _LIB = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _cm(a, dtype=np.float64):
    a = np.asarray(a, dtype=dtype)
    if a.ndim == 1:
        a = a[:, None]
    return np.asfortranarray(a)


def device_count() -> int:
    return load_library().globe_device_count()


REPLAY_PATH = os.path.join(HERE, "libglobe_plan_replay.so")


def plan_replay(digit_cols, m, n, r0=0, ncu=256, thin_bpc=(1,) * 8, shrink=(0, 0, 0)):
    """Host replay of both contraction passes of one product shape over every tuner
    candidate (libglobe_plan_replay.so; no GPU): raises AssertionError with the first
    read or store outside its buffer, or a partial slot left unwritten; returns the
    replay's counters. shrink: bytes taken off the packed layouts, the digit matrix and
    the partial buffer (to show that one access too many is caught)."""
    L = C.CDLL(REPLAY_PATH)
    f = L.globe_plan_replay
    f.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_char_p, C.c_int]
    f.restype = C.c_int
    bpc = np.ascontiguousarray(thin_bpc, dtype=np.int32)
    shr = np.ascontiguousarray(shrink, dtype=np.int64)
    st = np.zeros(8, dtype=np.int64)
    msg = C.create_string_buffer(512)
    rc = f(int(digit_cols), int(m), int(n), int(r0), int(ncu), _p(bpc), _p(shr), _p(st), msg, 512)
    assert rc == 0, msg.value.decode()
    return {"cand1": int(st[0]), "cand2": int(st[1]), "blocks": int(st[2]), "row_margin1": int(st[3]),
            "row_margin2": int(st[4]), "digit_bytes": int(st[5]), "part_bytes": int(st[6]), "wide": bool(st[7])}


def plan_replay_tg(m, n, ng, ncu=256, thin_bpc=(1,) * 8, shrink=(0, 0, 0)):
    """Host replay of the gathered pass 1 (k_geno_tg) over ng listed individuals of an
    n-individual cohort with m kept SNPs (libglobe_plan_replay.so; no GPU): raises
    AssertionError with the first read or store outside its buffer, LDS slot filled
    twice or read for the wrong piece, or SNP a split does not store exactly once;
    returns the launch (workgroups, splits), the row-read margin and the digit and
    partial buffers' bytes. shrink: bytes taken off the packed layout, the digit matrix
    and the partial buffer."""
    L = C.CDLL(REPLAY_PATH)
    f = L.globe_plan_replay_tg
    f.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p,
                  C.c_int]
    f.restype = C.c_int
    bpc = np.ascontiguousarray(thin_bpc, dtype=np.int32)
    shr = np.ascontiguousarray(shrink, dtype=np.int64)
    st = np.zeros(5, dtype=np.int64)
    msg = C.create_string_buffer(512)
    rc = f(int(m), int(n), int(ng), int(ncu), _p(bpc), _p(shr), _p(st), msg, 512)
    assert rc == 0, msg.value.decode()
    return {"workgroups": int(st[0]), "splits": int(st[1]), "row_margin": int(st[2]), "digit_bytes": int(st[3]),
            "part_bytes": int(st[4])}


class DeviceBuffer:
    """Device allocation owned by a context (for the *_dev entry points)."""

    def __init__(self, ctx: "Globe", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        ptr = C.c_void_p()
        ctx._check(ctx.L.globe_dev_alloc(ctx.h, self.nbytes, C.byref(ptr)))
        self.ptr = ptr.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr) if not arr.flags.f_contiguous else arr
        assert arr.nbytes <= self.nbytes
        self.ctx._check(self.ctx.L.globe_memcpy_h2d(self.ctx.h, self.ptr, _p(arr), arr.nbytes))

    def download(self, arr):
        assert arr.nbytes <= self.nbytes
        self.ctx._check(self.ctx.L.globe_memcpy_d2h(self.ctx.h, _p(arr), self.ptr, arr.nbytes))
        return arr

    def free(self):
        if self.ptr:
            self.ctx.L.globe_dev_free(self.ctx.h, self.ptr)
            self.ptr = None


class Globe:
    """One device context: cohort + resident genetics + variance components."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.globe_create(device, C.byref(h))
        if rc != GLOBE_OK:
            raise GlobeError(rc, f"globe_create(device={device}) failed "
                                 f"({self.L.globe_device_count()} device(s) visible)")
        self.h = h.value
        self.nrec = self.n = self.p = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.globe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != GLOBE_OK:
            msg = self.L.globe_last_error(self.h)
            raise GlobeError(rc, msg.decode() if msg else "")

    # ---- cohort / genetics (readInputData, prepGenotypeInput, readFullGRM)
    def set_cohort(self, indiv, times, X):
        self._indiv = np.ascontiguousarray(indiv, dtype=np.int32)
        self._times = np.ascontiguousarray(times, dtype=np.float64)
        X = _cm(X)
        self.nrec = int(self._indiv.shape[0])
        self.n = int(self._indiv.max()) + 1 if self.nrec else 0
        self.p = int(X.shape[1]) - 1
        self._check(self.L.globe_set_cohort(self.h, self.nrec, self.n, self.p, _p(self._indiv),
                                            _p(self._times), _p(X), X.shape[0]))

    def prepGenotypeInput(self, rows, min_maf=0.01, max_miss=0.2):
        rows = np.ascontiguousarray(rows, dtype=np.uint8)
        kept = C.c_int64()
        self._check(self.L.globe_load_bed(self.h, _p(rows), rows.shape[0], rows.shape[1],
                                          min_maf, max_miss, C.byref(kept)))
        self.m = kept.value
        self.m_total = rows.shape[0]
        return kept.value

    def load_bed_file(self, prefix, min_maf=0.01, max_miss=0.2):
        """prepGenotypeInput(prefix, ...) at file level: .bim line count, .bed header and
        size checks, .fam sample count when present, streamed ingest with device QC."""
        total, kept = C.c_int64(), C.c_int64()
        self._check(self.L.globe_load_bed_file(self.h, str(prefix).encode(), min_maf, max_miss,
                                               C.byref(total), C.byref(kept)))
        self.m = kept.value
        self.m_total = total.value
        return kept.value

    def set_ingest_chunk(self, nbytes):
        self._check(self.L.globe_set_ingest_chunk(self.h, int(nbytes)))

    def ingest_stats(self):
        """Last host-sourced ingest: total, host-fill and host-wait seconds, chunks, threads."""
        out = np.zeros(5)
        self._check(self.L.globe_ingest_stats(self.h, _p(out)))
        return {"seconds": out[0], "fill_s": out[1], "wait_s": out[2], "chunks": int(out[3]),
                "threads": int(out[4])}

    def synth_bed(self, m, seed, f_lo=0.01, f_hi=0.5, miss=0.0, min_maf=0.01, max_miss=0.2):
        kept = C.c_int64()
        self._check(self.L.globe_synth_bed(self.h, m, seed, f_lo, f_hi, miss, min_maf, max_miss,
                                           C.byref(kept)))
        self.m = kept.value
        self.m_total = m
        return kept.value

    def qc(self):
        m = self.m_total
        alt = np.zeros(m, dtype=np.int64)
        miss = np.zeros(m, dtype=np.int64)
        keep = np.zeros(m, dtype=np.uint8)
        maf = np.zeros(m, dtype=np.float32)
        self._check(self.L.globe_get_qc(self.h, _p(alt), _p(miss), _p(keep), _p(maf)))
        return {"alt": alt, "miss": miss, "keep": keep.astype(bool), "maf": maf}

    def allele_freqs(self):
        f = np.zeros(max(self.m, 1), dtype=np.float32)
        self._check(self.L.globe_get_freq(self.h, _p(f)))
        return f[: self.m]

    def geno_rows(self, first=0, count=None):
        count = self.m - first if count is None else count
        pitch = C.c_int64()
        self._check(self.L.globe_get_geno_rows(self.h, 0, 0, None, C.byref(pitch)))
        out = np.zeros((count, pitch.value), dtype=np.uint8)
        self._check(self.L.globe_get_geno_rows(self.h, first, count, _p(out), C.byref(pitch)))
        return out

    def readFullGRM(self, grm_lt):
        self._grm = np.ascontiguousarray(grm_lt, dtype=np.float32)
        self._check(self.L.globe_load_grm(self.h, _p(self._grm)))

    def computeGeneticDiagonal(self):
        out = np.zeros(self.n)
        self._check(self.L.globe_genetic_diag(self.h, _p(out)))
        return out

    def set_varcomp(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        self._check(self.L.globe_set_varcomp(self.h, _p(th)))

    # ---- operator (computeVx_multi) and preconditioner (applyPreconditioner)
    def computeVx_multi(self, x, multGRM_uhat=False, precision=PREC_FP64):
        x = _cm(x)
        k = x.shape[1]
        rows = self.n if multGRM_uhat else self.nrec
        y = np.zeros((rows, k), order="F")
        self._check(self.L.globe_vx(self.h, _p(x), x.shape[0], k, _p(y), rows,
                                    VX_GRM_ONLY if multGRM_uhat else VX_FULL, precision))
        return y

    def applyPreconditioner(self, r):
        r = _cm(r)
        z = np.zeros_like(r, order="F")
        self._check(self.L.globe_precond_apply(self.h, _p(r), r.shape[0], r.shape[1], _p(z),
                                               z.shape[0]))
        return z

    # ---- solver (runPCG_multi / runPCG_multi_VC)
    def pcg(self, b, *, tol=1e-5, max_iter=1000, x0_mode=X0_ZERO, rule=RULE_MAX_RATIO,
            stagnation=True, precision=PREC_FP64, x0=None, flags=0):
        b = _cm(b)
        k = b.shape[1]
        x = np.zeros_like(b, order="F") if x0 is None else _cm(x0).copy(order="F")
        o = PcgOpts(tol, max_iter, x0_mode, rule, int(stagnation), precision, flags)
        st = PcgStats()
        ratio = np.zeros(k)
        its = np.zeros(k, dtype=np.int32)
        self._check(self.L.globe_pcg(self.h, _p(b), b.shape[0], k, _p(x), x.shape[0],
                                     C.byref(o), C.byref(st), _p(ratio), _p(its)))
        return x, {"iters": st.iters, "stalled": st.stalled, "max_ratio": st.max_ratio,
                   "seconds": st.seconds, "col_ratio": ratio, "col_iters": its}

    def runPCG_multi(self, b, **kw):
        kw.setdefault("rule", RULE_MAX_RATIO)
        return self.pcg(b, **kw)

    def runPCG_multi_VC(self, b, **kw):
        kw.setdefault("rule", RULE_ALL_COLUMNS)
        kw.setdefault("stagnation", False)
        return self.pcg(b, **kw)

    # ---- estimators (runPCGsolver main body, runAIREML derivatives)
    def gls_blup(self, y, *, tol=1e-5, max_iter=1000, x0_mode=X0_ZERO, rule=RULE_MAX_RATIO,
                 stagnation=True, precision=PREC_FP64, flags=0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        beta = np.zeros(self.p + 1)
        vr = np.zeros(self.nrec)
        uh = np.zeros(self.n)
        bh = np.zeros(2 * self.n)
        o = PcgOpts(tol, max_iter, x0_mode, rule, int(stagnation), precision, flags)
        st = PcgStats()
        self._check(self.L.globe_gls_blup(self.h, _p(y), C.byref(o), _p(beta), _p(vr), _p(uh),
                                          _p(bh), C.byref(st)))
        return {"beta": beta, "vinv_r": vr, "uhat": uh, "bhat": bh, "iters": st.iters,
                "stalled": st.stalled, "seconds": st.seconds}

    def aireml_eval(self, y, probes, weights, probe_x0=None, *, tol=1e-5, max_iter=1000,
                    precision=PREC_FP64, reduced=False):
        """computeDerivative_first + estimateTrace_PdV + computeDerivative_AImat with
        injected probes (varCompEst.cpp:980-1041); reduced=True: the two-component model
        of varCompEst_reducedModel_noTime.cpp (:592-705)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        probes = _cm(probes)
        weights = _cm(weights)
        x0 = None if probe_x0 is None else _cm(probe_x0)
        nmc = probes.shape[1]
        nv = 2 if reduced else 5
        grad = np.zeros(nv)
        ai = np.zeros(nv * nv)
        first = np.zeros(nv)
        tr = np.zeros(nv)
        o = PcgOpts(tol, max_iter, X0_PRECOND, RULE_ALL_COLUMNS, 0, precision, 0)
        fn = self.L.globe_aireml_eval_reduced if reduced else self.L.globe_aireml_eval
        self._check(fn(self.h, _p(y), _p(probes), _p(weights), _p(x0), nmc,
                       C.byref(o), _p(grad), _p(ai), _p(first), _p(tr)))
        return {"grad": grad, "ai": ai.reshape(nv, nv), "first": first, "trace": tr}

    def aireml_probes(self, kind, nmc, seed):
        """The probes and trace weights the AI-REML fit draws on the device (Philox4x32-10
        keyed by seed): kind "gaussian" (tilde-V probes, weights tilde-V^-1 probe),
        "rademacher" or "exact"."""
        k = {"gaussian": 0, "rademacher": 1, "exact": 2}[kind]
        Y = np.zeros((self.nrec, nmc), order="F")
        Wt = np.zeros((self.nrec, nmc), order="F")
        self._check(self.L.globe_aireml_probes(self.h, k, nmc, C.c_uint64(seed), _p(Y), _p(Wt)))
        return Y, Wt

    def reml_loglik(self, y, *, probes="rademacher", nprobe=64, seed=1, tol=1e-8, max_iter=1000,
                    precision=PREC_FP64):
        """The REML log-likelihood at the current variance components: log|V| by
        stochastic Lanczos quadrature over `nprobe` probe solves ("gaussian" or
        "rademacher"), or exactly over all Nrec + 2N columns of the preconditioner's
        factor ("exact"); the GLS terms from the fit of gls_blup."""
        k = {"gaussian": 0, "rademacher": 1, "exact": 2}[probes]
        y = np.ascontiguousarray(y, dtype=np.float64)
        o = PcgOpts(tol, max_iter, X0_ZERO, RULE_ALL_COLUMNS, 0, precision, 0)
        r = Loglik()
        self._check(self.L.globe_reml_loglik(self.h, _p(y), k, int(nprobe), C.c_uint64(seed), C.byref(o),
                                             C.byref(r)))
        return {f: getattr(r, f) for f, _ in Loglik._fields_}

    def geno_Wa(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        out = np.zeros(self.n)
        self._check(self.L.globe_geno_Wa(self.h, _p(a), _p(out)))
        return out

    # ---- device-resident buffers
    def alloc(self, nbytes) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def pcg_dev(self, bbuf, xbuf, k, *, tol=1e-5, max_iter=1000, x0_mode=X0_ZERO,
                rule=RULE_PER_COLUMN, stagnation=False, precision=PREC_FP64, flags=0):
        o = PcgOpts(tol, max_iter, x0_mode, rule, int(stagnation), precision, flags)
        st = PcgStats()
        ratio = np.zeros(k)
        its = np.zeros(k, dtype=np.int32)
        self._check(self.L.globe_pcg_dev(self.h, bbuf.ptr, self.nrec, k, xbuf.ptr, self.nrec,
                                         C.byref(o), C.byref(st), _p(ratio), _p(its)))
        return {"iters": st.iters, "stalled": st.stalled, "max_ratio": st.max_ratio,
                "seconds": st.seconds, "col_ratio": ratio, "col_iters": its}

    def vx_dev(self, xbuf, ybuf, k, mode=VX_FULL, precision=PREC_FP64):
        rows = self.n if mode == VX_GRM_ONLY else self.nrec
        self._check(self.L.globe_vx_dev(self.h, xbuf.ptr, self.nrec, k, ybuf.ptr, rows, mode,
                                        precision))

    def synchronize(self):
        self._check(self.L.globe_synchronize(self.h))

    def set_mixed_digits(self, digits):
        """int8 digits (2..4) of mixed-precision genetic products outside a PCG solve."""
        self._check(self.L.globe_set_mixed_digits(self.h, int(digits)))

    def set_graphs(self, on=True):
        """PCG iterations replayed from captured HIP graphs (on) or enqueued kernel by
        kernel (off, the default); the results are identical."""
        self._check(self.L.globe_set_graphs(self.h, int(on)))

    def set_row_gather(self, on=True):
        """Serving loop: form only the requests' individuals' rows in the genotype
        product's individual-major pass (on, the default) or every row (off)."""
        self._check(self.L.globe_set_row_gather(self.h, int(on)))

    def device_memory(self):
        """(free, total) bytes of device memory."""
        f, t = C.c_int64(), C.c_int64()
        self._check(self.L.globe_device_memory(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    # ---- contraction variants
    def force_variant(self, index):
        """index >= 0: every genotype contraction runs candidate index % count of its
        shape's tuner candidates (untimed); -1 returns to the tuned variants."""
        self._check(self.L.globe_force_variant(self.h, int(index)))

    def variant_counts(self):
        """Candidates of pass 1 and pass 2 of the last genotype contraction."""
        out = np.zeros(2, dtype=np.int32)
        self._check(self.L.globe_variant_counts(self.h, _p(out)))
        return int(out[0]), int(out[1])

    # ---- kernel timing
    def profile(self, on=True):
        self._check(self.L.globe_profile_enable(self.h, int(on)))

    def kernel_stats(self, which):
        st = KernelStats()
        self._check(self.L.globe_profile_read(self.h, which, C.byref(st)))
        return {"launches": st.launches, "ms_total": st.ms_total,
                "ms_avg": st.ms_total / st.launches if st.launches else 0.0,
                "bytes": st.bytes_per_launch, "flops": st.flops_per_launch,
                "issued": st.issued_ops_per_launch}

    def grm_from_geno(self):
        """Build G = W W^T / M from the resident genotypes; the context then uses the GRM."""
        self._check(self.L.globe_grm_from_geno(self.h))

    def grm_rows(self, r0, r1):
        """Rows r0 .. r1 - 1 of the GRM's packed lower triangle (.grm.bin order)."""
        out = np.zeros(r1 * (r1 + 1) // 2 - r0 * (r0 + 1) // 2, dtype=np.float32)
        self._check(self.L.globe_get_grm_rows(self.h, int(r0), int(r1), _p(out)))
        return out

    def grm_lt(self):
        lt = np.zeros(self.n * (self.n + 1) // 2, dtype=np.float32)
        self._check(self.L.globe_get_grm(self.h, _p(lt)))
        return lt

    # ---- serving loop
    def serve_open(self, *, slots=32, max_records, hbm_budget=0, tol=1e-6, max_iter=1000,
                   precision=PREC_MIXED):
        o = ServeOpts(slots, max_records, hbm_budget, tol, max_iter, precision)
        self._check(self.L.globe_serve_open(self.h, C.byref(o)))
        self._serve_nrec = {}

    def serve_submit(self, indiv, times, y, theta):
        """Queue a solve V x = y for the records (indiv, times); returns the request id."""
        indiv = np.ascontiguousarray(indiv, dtype=np.int32)
        times = np.ascontiguousarray(times, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        r = Request(indiv.shape[0], indiv.ctypes.data, times.ctypes.data, y.ctypes.data,
                    (C.c_double * 5)(*theta))
        rid = C.c_int64()
        # the request's id is recorded before the call returns to the caller, so a stepping
        # thread can take it as soon as it is finished
        self._check(self.L.globe_serve_submit(self.h, C.byref(r), C.byref(rid)))
        self._serve_nrec[rid.value] = indiv.shape[0]
        return rid.value

    def serve_step(self):
        """One admission + block-PCG iteration; returns (finished, active, queued)."""
        f, a, q = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.L.globe_serve_step(self.h, C.byref(f), C.byref(a), C.byref(q)))
        return f.value, a.value, q.value

    def serve_finished(self, max_ids=4096):
        ids = np.zeros(max_ids, dtype=np.int64)
        n = C.c_int32()
        self._check(self.L.globe_serve_finished(self.h, _p(ids), max_ids, C.byref(n)))
        return [int(v) for v in ids[:n.value]]

    def serve_take(self, rid):
        nrec = self._serve_nrec.get(rid)
        if nrec is None:
            raise KeyError(f"request {rid}: serve_submit has not returned it to this handle")
        x = np.zeros(nrec)
        res = ServeResult()
        self._check(self.L.globe_serve_take(self.h, rid, _p(x), C.byref(res)))
        del self._serve_nrec[rid]
        return x, {"id": res.id, "iters": res.iters, "converged": bool(res.converged), "ratio": res.ratio,
                   "t_submit": res.t_submit, "t_admit": res.t_admit, "t_done": res.t_done}

    def serve_usage(self):
        """Record pool: bytes in use, peak bytes in use, pool bytes, steps the queue's head
        waited for room."""
        out = np.zeros(4, dtype=np.int64)
        self._check(self.L.globe_serve_usage(self.h, _p(out)))
        return {"in_use": int(out[0]), "peak": int(out[1]), "pool": int(out[2]), "waits": int(out[3])}

    def serve_close(self):
        self._check(self.L.globe_serve_close(self.h))

    # ---- AI-REML (runAIREML)
    def aireml_fit(self, y, theta0, *, max_iter=50, tol=1e-3, num_mc=15, probes="gaussian",
                   seed=20261203, pcg_tol=1e-5, precision=PREC_FP64, log=None, reduced=False):
        """runAIREML (varCompEst.cpp:1347-1500): theta0 has five components; reduced=True:
        runAIREML_reducedModel (varCompEst_reducedModel_noTime.cpp:967-1099), theta0 =
        (sig2e, sig2g). probes: "gaussian" (the reference's tilde-V probes), "rademacher"
        or "exact" (unit probes; num_mc must equal the record count)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        kind = {"gaussian": 0, "rademacher": 1, "exact": 2}[probes]
        o = AiremlOpts(max_iter, tol, num_mc, kind, seed,
                       PcgOpts(pcg_tol, 1000, X0_PRECOND, RULE_ALL_COLUMNS, 0, precision, 0))
        nv = 2 if reduced else 5
        th = np.ascontiguousarray(theta0, dtype=np.float64).copy()
        if th.shape != (nv,):
            raise ValueError(f"theta0 must have {nv} components")
        se = np.zeros(nv)
        ai = np.zeros(nv * nv)
        it = C.c_int32()
        spi = C.c_double()
        lines = []
        cb = LOG_FN(lambda s, u: (lines.append(s.decode()), log(s.decode()) if log else None))
        fn = self.L.globe_aireml_fit_reduced if reduced else self.L.globe_aireml_fit
        self._check(fn(self.h, _p(y), C.byref(o), _p(th), _p(se), _p(ai),
                       C.byref(it), C.byref(spi), cb, None))
        return {"theta": th, "se": se, "ai": ai.reshape(nv, nv), "iters": it.value,
                "sec_per_iter": spi.value, "log": lines}
