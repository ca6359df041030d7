"""ORACLE (test infrastructure only): ctypes binding of the C restatement in
oracle/hc_oracle.c.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package, and only as the checker / CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
sys.path.insert(0, str(REPO / "homotopycontinuation.jl_amd"))

from hcgpu.abi import (  # noqa: E402  (struct layouts shared with the C ABI)
    RETURN_CODES, TRACKER_CODES, default_options, hcg_batch, hcg_homotopy_desc, hcg_options, hcg_path_result,
)
from hcgpu.tape import hcg_tape  # noqa: E402

LIB = HERE / "_build" / "liboracle.so"
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def use_native_build() -> str:
    """Switch this process to the timing build of the same source (-O3 -march=native, made on
    this host; same arithmetic contract, so identical results).  Must run before lib()."""
    global LIB
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    subprocess.run(["make", "-s", "-B", "-C", str(HERE), "native"], check=True)  # always for this host
    LIB = HERE / "_build" / "liboracle_native.so"
    return "gcc -O3 -march=native -ffp-contract=off"


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        P = ctypes.POINTER
        D = ctypes.c_double
        L.orc_track_batch.argtypes = [P(hcg_homotopy_desc), P(hcg_batch), P(hcg_options), ctypes.c_int,
                                      P(hcg_path_result), P(D), P(D), P(D)]
        L.orc_track_batch_log.argtypes = [P(hcg_homotopy_desc), P(hcg_batch), P(hcg_options), ctypes.c_int,
                                          P(hcg_path_result), P(D), P(D), P(D), P(ctypes.c_float), ctypes.c_int,
                                          P(ctypes.c_int)]
        L.orc_track_plain.argtypes = [P(hcg_homotopy_desc), P(hcg_options), P(D), P(D), P(D), D, D, P(D),
                                      P(ctypes.c_int32), P(D)]
        L.orc_predictor_probe.argtypes = [P(hcg_homotopy_desc), P(hcg_options), P(D), P(D), ctypes.c_int,
                                          ctypes.c_int, P(D), P(D), P(D), D, D, P(D), P(D), P(D), P(D)]
        L.orc_sys_eval.argtypes = [P(hcg_tape), P(D), P(D), ctypes.c_int, P(D), P(D)]
        L.orc_sys_taylor.argtypes = [P(hcg_tape), ctypes.c_int, P(D), P(D), P(D), ctypes.c_int, P(D)]
        L.orc_sys_eval_dd.argtypes = [P(hcg_tape), P(D), P(D), P(D), ctypes.c_int, P(D), P(D)]
        L.orc_dd_op.argtypes = [ctypes.c_int, D, D, D, D, P(D), P(D)]
        L.orc_dd_op.restype = None
        L.orc_linsolve.argtypes = [ctypes.c_int, P(D), P(D), P(D), P(D), P(D), P(D), P(D)]
        L.orc_rank.argtypes = [ctypes.c_int, P(D), D]
        L.orc_norm.argtypes = [ctypes.c_int, P(D), P(D)]
        L.orc_norm.restype = D
        L.orc_nthroot.argtypes = [D, ctypes.c_int]
        L.orc_nthroot.restype = D
        L.orc_eps.argtypes = [D]
        L.orc_eps.restype = D
        L.orc_valuation.argtypes = [ctypes.c_int, P(D), D, P(D), P(D)]
        L.orc_valuation.restype = None
        L.orc_sizeof.argtypes = [ctypes.c_int]
        L.orc_sizeof.restype = ctypes.c_size_t
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_certify.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, P(D), ctypes.c_int, P(D), P(D), P(D),
                                  P(ctypes.c_int32), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_void_p, P(D), P(D), P(D)]
        L.orc_ival_op.argtypes = [ctypes.c_int, P(D), P(D), P(D), P(D)]
        L.orc_ival_op.restype = None
        L.orc_ball_op.argtypes = [ctypes.c_int, P(D), P(D), P(D)]
        L.orc_ball_op.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


class BatchResult:
    """Per-path arrays of one batch (PathResult fields)."""

    def __init__(self, res, sol, last, val):
        self.raw = res
        self.code = np.array([r.code for r in res], dtype=np.int32)
        self.singular = np.array([r.singular for r in res], dtype=bool)
        self.winding_number = np.array([r.winding_number for r in res], dtype=np.int32)
        self.accepted_steps = np.array([r.accepted_steps for r in res], dtype=np.int64)
        self.rejected_steps = np.array([r.rejected_steps for r in res], dtype=np.int64)
        self.steps_eg = np.array([r.steps_eg for r in res], dtype=np.int64)
        self.newton_iters = np.array([r.newton_iters for r in res], dtype=np.int64)
        self.extended_precision = np.array([r.extended_precision for r in res], dtype=bool)
        self.extended_precision_used = np.array([r.extended_precision_used for r in res], dtype=bool)
        self.valuation_valid = np.array([r.valuation_valid for r in res], dtype=bool)
        self.t = np.array([r.t for r in res])
        self.accuracy = np.array([r.accuracy for r in res])
        self.residual = np.array([r.residual for r in res])
        self.cond = np.array([r.cond for r in res])
        self.omega = np.array([r.omega for r in res])
        self.mu = np.array([r.mu for r in res])
        self.last_t = np.array([r.last_t for r in res])
        self.solution = sol
        self.last_point = last
        self.valuation = val

    def codes(self):
        return [RETURN_CODES[c] for c in self.code]


def track_batch(H, starts, opts: hcg_options | None = None, nthreads: int = 0, group_of_path=None,
                log_capacity: int = 0) -> BatchResult:
    """Track every start with the oracle (threaded_solve restatement)."""
    opts = opts or default_options()
    starts = _c(starts)
    npaths, n = starts.shape
    b = hcg_batch()
    b.npaths = npaths
    b.starts = _dp(starts)
    keep = [starts]
    if getattr(H, "q", None) is not None:
        p = _c(H.p)
        q = _c(H.q)
        keep += [p, q]
        b.start_params = _dp(p)
        b.target_params = _dp(q)
        b.ngroups = q.shape[0]
        if group_of_path is not None:
            g = np.ascontiguousarray(np.asarray(group_of_path, dtype=np.int32))
            keep.append(g)
            b.group_of_path = g.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    b.t1, b.t0 = 1.0, 0.0
    res = (hcg_path_result * npaths)()
    sol = np.zeros((npaths, n), dtype=np.complex128)
    last = np.zeros((npaths, n), dtype=np.complex128)
    val = np.zeros((npaths, n), dtype=np.float64)
    nthreads = nthreads or len(os.sched_getaffinity(0))
    log = np.zeros((max(log_capacity, 1), 16), dtype=np.float32)
    cnt = ctypes.c_int(0)
    rc = lib().orc_track_batch_log(ctypes.byref(H.desc()), ctypes.byref(b), ctypes.byref(opts), nthreads, res,
                                   _dp(sol), _dp(last), _dp(val),
                                   log.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if log_capacity > 0
                                   else ctypes.POINTER(ctypes.c_float)(), log_capacity, ctypes.byref(cnt))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    out = BatchResult(list(res), sol, last, val)
    out.log = log[: min(cnt.value, log_capacity)] if log_capacity > 0 else None
    return out


def track_plain(H, x1, t1=1.0, t0=0.0, opts: hcg_options | None = None, p=None, q=None):
    """Tracker-level track(tracker, x, t₁, t₀) -> dict (TrackerResult fields)."""
    opts = opts or default_options()
    x1 = _c(x1)
    n = x1.shape[0]
    xo = np.zeros(n, dtype=np.complex128)
    io = (ctypes.c_int32 * 5)()
    do = np.zeros(6)
    if p is None and hasattr(H, "p"):
        p, q = H.p, H.q[0]
    pp = _c(p) if p is not None else np.zeros(1, dtype=np.complex128)
    qq = _c(q) if q is not None else np.zeros(1, dtype=np.complex128)
    rc = lib().orc_track_plain(ctypes.byref(H.desc()), ctypes.byref(opts), _dp(pp), _dp(qq), _dp(x1), t1, t0,
                               _dp(xo), io, _dp(do))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return dict(return_code=TRACKER_CODES[io[0]], solution=xo, accepted_steps=io[1], rejected_steps=io[2],
                extended_precision=bool(io[3]), extended_precision_used=bool(io[4]), t=complex(do[0], do[1]),
                accuracy=do[2], omega=do[3], mu=do[4], tau=do[5])


def predictor_probe(H, xs, ts, dt, winding=1, w=None, p=None, q=None, opts: hcg_options | None = None):
    """The tracker's predictor on given path points (orc_predictor_probe): predict! from each
    point to the next, update! at every point, then predict! by dt from the last one.
    Returns dict(tx = 4 x n Taylor coefficients x⁰..x³ after the last update, xh = the final
    prediction, xhats = the prediction made before each update (row 0 unused), and per update
    trust_region, local_error, cond_H_xd, hermite, order)."""
    opts = opts or default_options()
    xs = np.ascontiguousarray(np.asarray(xs, dtype=np.complex128).reshape(len(ts), -1))
    npts, n = xs.shape
    tsc = _c(np.asarray(ts, dtype=np.complex128))
    if p is None and hasattr(H, "p"):
        p, q = H.p, H.q[0]
    pp = _c(p) if p is not None else np.zeros(1, dtype=np.complex128)
    qq = _c(q) if q is not None else np.zeros(1, dtype=np.complex128)
    ww = np.ascontiguousarray(np.asarray(w, dtype=np.float64)) if w is not None else None
    tx = np.zeros(4 * n, dtype=np.complex128)
    xh = np.zeros(n, dtype=np.complex128)
    xhats = np.zeros(npts * n, dtype=np.complex128)
    scal = np.zeros(5 * npts)
    dt = complex(dt)
    rc = lib().orc_predictor_probe(ctypes.byref(H.desc()), ctypes.byref(opts), _dp(pp), _dp(qq), int(winding), npts,
                                   _dp(xs), _dp(tsc), _dp(ww) if ww is not None else ctypes.POINTER(ctypes.c_double)(),
                                   dt.real, dt.imag, _dp(tx), _dp(xh), _dp(xhats), _dp(scal))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    sc = scal.reshape(npts, 5)
    return dict(tx=tx.reshape(4, n), xh=xh, xhats=xhats.reshape(npts, n), trust_region=sc[:, 0],
                local_error=sc[:, 1], cond_H_xd=sc[:, 2], hermite=sc[:, 3].astype(int), order=sc[:, 4].astype(int))


def sys_eval(tape, x, p=(), n=None):
    x = _c(x)
    n = n or x.shape[0]
    p = _c(p) if len(p) else np.zeros(1, dtype=np.complex128)
    u = np.zeros(tape.output_dim, dtype=np.complex128)
    U = np.zeros(n * n, dtype=np.complex128)
    rc = lib().orc_sys_eval(ctypes.byref(tape.struct()), _dp(x), _dp(p), n, _dp(u), _dp(U))
    assert rc == 0, lib().orc_last_error()
    return u, U.reshape(n, n).T  # column-major -> [i, j]


def sys_taylor(tape, K, tx, p0=None, p1=None):
    tx = _c(tx).reshape(K, -1)
    n = tx.shape[1]
    out = np.zeros((K + 1) * tape.output_dim, dtype=np.complex128)
    q0 = _c(p0) if p0 is not None else None
    q1 = _c(p1) if p1 is not None else None
    null = ctypes.POINTER(ctypes.c_double)()
    rc = lib().orc_sys_taylor(ctypes.byref(tape.struct()), K, _dp(tx), _dp(q0) if q0 is not None else null,
                              _dp(q1) if q1 is not None else null, n, _dp(out))
    assert rc == 0, lib().orc_last_error()
    return out.reshape(K + 1, tape.output_dim)


def sys_eval_dd(tape, xhi, xlo, p=()):
    xhi, xlo = _c(xhi), _c(xlo)
    n = xhi.shape[0]
    p = _c(p) if len(p) else np.zeros(1, dtype=np.complex128)
    uh = np.zeros(tape.output_dim, dtype=np.complex128)
    ul = np.zeros(tape.output_dim, dtype=np.complex128)
    rc = lib().orc_sys_eval_dd(ctypes.byref(tape.struct()), _dp(xhi), _dp(xlo), _dp(p), n, _dp(uh), _dp(ul))
    assert rc == 0
    return uh, ul


def dd_op(op, a, b):
    rh, rl = ctypes.c_double(), ctypes.c_double()
    lib().orc_dd_op(op, a[0], a[1], b[0], b[1], ctypes.byref(rh), ctypes.byref(rl))
    return rh.value, rl.value


def linsolve(A, b, weights=None, dl=None, dr=None):
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    Ac = _c(A.T.reshape(-1))  # column-major
    bb = _c(b)
    x = np.zeros(n, dtype=np.complex128)
    cond = ctypes.c_double()
    null = ctypes.POINTER(ctypes.c_double)()
    w = np.ascontiguousarray(weights, dtype=np.float64) if weights is not None else None
    l = np.ascontiguousarray(dl, dtype=np.float64) if dl is not None else None
    r = np.ascontiguousarray(dr, dtype=np.float64) if dr is not None else None
    lib().orc_linsolve(n, _dp(Ac), _dp(bb), _dp(w) if w is not None else null, _dp(x), ctypes.pointer(cond),
                       _dp(l) if l is not None else null, _dp(r) if r is not None else null)
    return x, cond.value


def rank(A, rtol=1e-14):
    A = np.asarray(A, dtype=np.complex128)
    return lib().orc_rank(A.shape[0], _dp(_c(A.T.reshape(-1))), rtol)


def norm(x, w=None):
    x = _c(x)
    null = ctypes.POINTER(ctypes.c_double)()
    ww = np.ascontiguousarray(w, dtype=np.float64) if w is not None else None
    return lib().orc_norm(x.shape[0], _dp(x), _dp(ww) if ww is not None else null)


def track_plain_valuation(H, x1, t1, t0, opts: hcg_options | None = None):
    opts = opts or default_options()
    L = lib()
    L.orc_track_plain_valuation.argtypes = [ctypes.POINTER(hcg_homotopy_desc), ctypes.POINTER(hcg_options),
                                            ctypes.POINTER(ctypes.c_double), ctypes.c_double, ctypes.c_double,
                                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    x1 = _c(x1)
    n = x1.shape[0]
    vx = np.zeros(n)
    vt = np.zeros(n)
    code = L.orc_track_plain_valuation(ctypes.byref(H.desc()), ctypes.byref(opts), _dp(x1), t1, t0, _dp(vx), _dp(vt))
    return TRACKER_CODES[code], vx, vt


def math_fn(which, x, y=0.0):
    L = lib()
    L.orc_math.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.orc_math.restype = ctypes.c_double
    return L.orc_math(which, x, y)


def segment_run(start, target, ops):
    """Run a SegmentStepper script (propose_step! for ops >= 0, step_success! for -1); returns one
    dict per op with s, s1, ds, t, t1, dt, abs_delta, forward, done, dist."""
    L = lib()
    L.orc_segment_run.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                                           ctypes.POINTER(ctypes.c_double)]
    start, target = complex(start), complex(target)
    o = np.ascontiguousarray(ops, dtype=np.float64)
    out = np.zeros(14 * len(o))
    rc = L.orc_segment_run(start.real, start.imag, target.real, target.imag, len(o), _dp(o), _dp(out))
    assert rc == 0
    rows = []
    for r in out.reshape(-1, 14):
        rows.append(dict(s=r[0], s1=r[1], ds=r[2], t=complex(r[3], r[4]), t1=complex(r[5], r[6]),
                         dt=complex(r[7], r[8]), abs_delta=r[9], forward=bool(r[10]), done=bool(r[11]), dist=r[12]))
    return rows


class CertifyOutput:
    """Per-candidate arrays of one certification batch (hcg_certificate fields, the boxes
    x₁ and x₀ as (nsol, n, 4) arrays of (re lo, re hi, im lo, im hi), refined candidates)."""

    def __init__(self, rec, box1, box0, refined):
        self.records = rec
        for f in rec.dtype.names:
            setattr(self, f, rec[f])
        self.box1, self.box0, self.refined = box1, box0, refined


def certify(system, n, candidates, params=None, params_ival=None, params_ball=None, refine=True, extended=True,
            real_system=True, nthreads=0, group_of=None):
    """The oracle's certification of `candidates` (nsol x n) for an hcg_system struct.
    params / params_ival / params_ball: one group (nparams, ...) or several (ngroups, nparams,
    ...); group_of: the group of each candidate (default group 0)."""
    from hcgpu.abi import hcg_certificate
    X = _c(candidates).reshape(-1, n)
    nsol = len(X)
    rec = np.zeros(nsol, dtype=np.dtype([(f, np.int32 if t is ctypes.c_int32 else np.float64)
                                         for f, t in hcg_certificate._fields_], align=True))
    box1 = np.zeros((nsol, n, 4))
    box0 = np.zeros((nsol, n, 4))
    ref = np.zeros((nsol, n), dtype=np.complex128)
    P = _c(params) if params is not None else None
    ngroups = 0 if P is None else (P.shape[0] if P.ndim == 2 else 1)
    IP = np.ascontiguousarray(np.asarray(params_ival, dtype=np.float64)) if params_ival is not None else None
    BP = np.ascontiguousarray(np.asarray(params_ball, dtype=np.float64)) if params_ball is not None else None
    G = np.ascontiguousarray(np.asarray(group_of, dtype=np.int32)) if group_of is not None else None
    keep = (X, P, IP, BP, G)  # noqa: F841  (alive during the call)
    nthreads = nthreads or min(len(os.sched_getaffinity(0)), 16)
    opt = lambda a: _dp(a) if a is not None else None  # noqa: E731
    rc = lib().orc_certify(ctypes.cast(ctypes.pointer(system), ctypes.c_void_p), n, nsol, _dp(X), ngroups, opt(P),
                           opt(IP), opt(BP), G.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if G is not None else None,
                           int(refine), int(extended), int(real_system), nthreads,
                           rec.ctypes.data_as(ctypes.c_void_p), _dp(box1), _dp(box0), _dp(ref))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return CertifyOutput(rec, box1, box0, ref)


def ival_op(op, a, b=(0.0, 0.0), c=(0.0, 0.0)):
    """One interval operation of the oracle (see orc_ival_op): returns (lo, hi)."""
    A, B, Cc = (np.asarray(v, dtype=np.float64) for v in (a, b, c))
    out = np.zeros(2)
    lib().orc_ival_op(op, _dp(A), _dp(B), _dp(Cc), _dp(out))
    return float(out[0]), float(out[1])


def ball_op(op, a, b):
    """One double-double ball operation of the oracle (0 add, 1 sub, 2 mul) on balls given as
    (hi, lo, radius); returns (hi, lo, radius)."""
    A, B = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_ball_op(op, _dp(A), _dp(B), _dp(out))
    return tuple(float(v) for v in out)
