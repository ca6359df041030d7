"""Host-side handles of the C ABI: GPU context, compiled programs, batch tracking."""
from __future__ import annotations

import ctypes
import threading
from typing import Dict, Optional

import numpy as np

from . import abi
from .abi import check, hcg_batch, hcg_options, hcg_path_result, hcg_step_log, lib


class GpuContext:
    """One hcg_ctx (device + stream).  Not thread-safe: one per host thread."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        check(lib().hcg_ctx_create(device, ctypes.byref(self._h)))
        self.device = device
        self._programs: Dict[int, "Program"] = {}
        self._tuned: Dict[int, dict] = {}  # autotuner choices looked up per homotopy

    @property
    def handle(self):
        return self._h

    def program(self, homotopy, waves_per_eu: int = 0, layout: Optional[str] = None) -> "Program":
        """layout "lane": one path per lane, the tapes compiled with hipRTC; "group": one path
        per group of 16 lanes on the precompiled lane-group kernel (systems of up to 16
        variables; 32 lanes for 17 to 24 variables).  With neither layout nor waves_per_eu
        given, the autotuner's cached choice for this homotopy on this device
        (hcgpu.autotune.tune) when there is one, including its workgroups per CU, which
        track_batch then uses; otherwise one path per lane at the library's default register
        budget up to LANE_MAX_N variables, and the lane-group layout past that: there the
        per-lane kernel's LU leaves the registers for private memory (measured on 8,192-path
        batches: 4.6x faster at n = 15, 8x at n = 24)."""
        blocks_per_cu = 0
        tuned = None
        if layout is None and waves_per_eu == 0:
            tuned = self._tuned.get(id(homotopy))
            if tuned is None:
                from . import autotune
                tuned = autotune.tuned_choice(self, homotopy) or {}
                self._tuned[id(homotopy)] = tuned
            if tuned:
                layout, waves_per_eu, blocks_per_cu = tuned["layout"], tuned["waves_per_eu"], tuned["blocks_per_cu"]
        layout = layout or ("group" if homotopy.n > LANE_MAX_N and waves_per_eu == 0 else "lane")
        key = (id(homotopy), waves_per_eu, layout)
        p = self._programs.get(key)
        if p is None:
            p = Program(self, homotopy, waves_per_eu, layout)
            self._programs[key] = p
        if tuned:
            p.tuned = dict(tuned)
            p.blocks_per_cu = blocks_per_cu
        return p

    def program_multi(self, homotopies, waves_per_eu: int = 0, layout: str = "lane") -> "MultiProgram":
        key = (tuple(id(h) for h in homotopies), waves_per_eu, layout)
        p = self._programs.get(key)
        if p is None:
            p = MultiProgram(self, list(homotopies), waves_per_eu, layout)
            self._programs[key] = p
        return p

    def certifier(self, system, fixed_params=None) -> "Certifier":
        key = ("certifier", id(system), None if fixed_params is None else tuple(np.asarray(fixed_params).ravel()))
        c = self._programs.get(key)
        if c is None:
            c = Certifier(self, system, fixed_params)
            self._programs[key] = c
        return c

    def last_kernel_ms(self) -> float:
        ms = ctypes.c_double()
        check(lib().hcg_last_kernel_stats(self._h, ctypes.byref(ms), None, None))
        return ms.value

    def last_counters(self):
        out = (ctypes.c_uint64 * 4)()
        check(lib().hcg_last_counters(self._h, out))
        return dict(steps=int(out[0]), newton_iters=int(out[1]), flops=int(out[2]), learned_accepted=int(out[3]))

    def close(self):
        for p in self._programs.values():
            p.close()
        self._programs.clear()
        if self._h:
            lib().hcg_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[GpuContext] = None
_lock = threading.Lock()


def default_context() -> GpuContext:
    global _default_ctx
    with _lock:
        if _default_ctx is None:
            _default_ctx = GpuContext(0)
        return _default_ctx


LAYOUTS = {"lane": 0, "group": 1}
# largest system the per-lane kernel factors in registers (its left-looking register LU)
LANE_MAX_N = 12


def group_schedule_info(homotopies):
    """Slot schedules of the lane-group layout for these homotopies (no device needed): the
    kernel's size class, LDS bytes per path and, per homotopy, slots / registers /
    instructions of its eval_jac and eval programs."""
    descs = [h.desc() for h in homotopies]
    arr = (ctypes.POINTER(abi.hcg_homotopy_desc) * len(descs))(*[ctypes.pointer(d) for d in descs])
    out = (ctypes.c_int32 * (2 + 6 * len(descs)))()
    check(lib().hcg_group_schedule_info(ctypes.cast(arr, ctypes.c_void_p), len(descs), out))
    systems = []
    for s in range(len(descs)):
        o = out[2 + 6 * s: 8 + 6 * s]
        systems.append(dict(ej_slots=o[0], ej_regs=o[1], ej_instr=o[2], e_slots=o[3], e_regs=o[4], e_instr=o[5]))
    return dict(nn=out[0], lds_per_path=out[1], systems=systems)


def _create(ctx, descs, multi: bool, waves_per_eu: int, layout: str = "lane"):
    if layout not in LAYOUTS:
        raise ValueError(f"layout must be one of {sorted(LAYOUTS)}")
    h = ctypes.c_void_p()
    arr = (ctypes.POINTER(abi.hcg_homotopy_desc) * len(descs))(*[ctypes.pointer(d) for d in descs])
    bo = abi.hcg_build_options()
    bo.waves_per_eu = waves_per_eu
    bo.layout = LAYOUTS[layout]
    L = lib()
    L.hcg_program_create_ex.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(abi.hcg_homotopy_desc)),
                                        ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(abi.hcg_build_options),
                                        ctypes.POINTER(ctypes.c_void_p)]
    check(L.hcg_program_create_ex(ctx.handle, arr, len(descs), int(multi), ctypes.byref(bo), ctypes.byref(h)))
    return h


class Program:
    """A homotopy lowered to HIP and compiled with hipRTC (hcg_program_create_ex);
    waves_per_eu = 0 takes the library's default register budget."""

    def __init__(self, ctx: GpuContext, homotopy, waves_per_eu: int = 0, layout: str = "lane"):
        self.ctx = ctx
        self.homotopy = homotopy
        self.n = homotopy.n
        self.waves_per_eu = waves_per_eu
        self.layout = layout
        self.blocks_per_cu = 0  # launch shape from the autotuner (0: the library's default)
        self.tuned = None
        self._h = _create(ctx, [homotopy.desc()], False, waves_per_eu, layout)

    def source(self) -> str:
        return lib().hcg_program_source(self._h).decode()

    def kernel_info(self):
        return _kernel_info(self._h)

    def close(self):
        if self._h:
            lib().hcg_program_destroy(self._h)
            self._h = ctypes.c_void_p()


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class BatchOutput:
    """Per-path arrays of one hcg_track_batch call."""

    FIELDS = [f for f, _ in hcg_path_result._fields_]

    def __init__(self, rec: np.ndarray, sol, last, val, log=None):
        self.records = rec
        for f in self.FIELDS:
            setattr(self, f, rec[f])
        self.singular = self.singular.astype(bool)
        self.solution = sol
        self.last_point = last
        self.valuation = val
        self.log = log

    def codes(self):
        return [abi.RETURN_CODES[c] for c in self.code]


_REC_DTYPE = np.dtype([(f, np.int32 if t is ctypes.c_int32 else np.float64) for f, t in hcg_path_result._fields_],
                      align=True)
assert _REC_DTYPE.itemsize == ctypes.sizeof(hcg_path_result)


def track_batch(program: Program, starts, opts: Optional[hcg_options] = None, target_groups=None,
                group_of_path=None, log_capacity: int = 0, start_params=None) -> BatchOutput:
    """Track all `starts` (npaths x n complex) from t=1 to t=0 on the GPU.  For a parameter
    homotopy, `start_params` / `target_groups` replace the homotopy's own p / q (in the
    space its tapes use, see ParameterHomotopy.target_matrix)."""
    opts = hcg_options.from_buffer_copy(opts) if opts is not None else abi.default_options()
    if opts.blocks_per_cu == 0 and getattr(program, "blocks_per_cu", 0):
        opts.blocks_per_cu = program.blocks_per_cu  # the autotuner's launch shape
    starts = np.ascontiguousarray(np.asarray(starts, dtype=np.complex128))
    npaths, n = starts.shape
    if n != program.n:
        raise ValueError(f"start solutions have {n} coordinates, the homotopy has {program.n}")
    b = hcg_batch()
    b.npaths = npaths
    b.starts = _dp(starts)
    b.t1, b.t0 = 1.0, 0.0
    keep = [starts]
    H = program.homotopy
    if getattr(H, "q", None) is not None:
        p = np.ascontiguousarray(H.p if start_params is None else np.asarray(start_params, dtype=np.complex128))
        if p.shape != np.shape(H.p):
            raise ValueError("start_params has the wrong length")
        q = np.ascontiguousarray(H.q if target_groups is None else np.asarray(target_groups, dtype=np.complex128))
        keep += [p, q]
        b.start_params, b.target_params = _dp(p), _dp(q)
        b.ngroups = q.shape[0]
        if group_of_path is not None:
            g = np.ascontiguousarray(np.asarray(group_of_path, dtype=np.int32))
            if g.shape != (npaths,):
                raise ValueError("group_of_path must have one entry per path")
            keep.append(g)
            b.group_of_path = g.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    rec = np.zeros(npaths, dtype=_REC_DTYPE)
    sol = np.zeros((npaths, n), dtype=np.complex128)
    last = np.zeros((npaths, n), dtype=np.complex128)
    val = np.zeros((npaths, n), dtype=np.float64)
    logp = None
    log_arr = None
    if log_capacity > 0:
        opts.log_steps = 1
        log_arr = np.zeros((log_capacity, abi.HCG_STEP_RECORD_FLOATS), dtype=np.float32)
        cnt = ctypes.c_int32(0)
        lg = hcg_step_log()
        lg.capacity = log_capacity
        lg.records = log_arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        lg.count = ctypes.pointer(cnt)
        logp = ctypes.byref(lg)
        keep += [log_arr, cnt, lg]
    check(lib().hcg_track_batch(program.ctx.handle, program._h, ctypes.byref(b), ctypes.byref(opts),
                                rec.ctypes.data_as(ctypes.POINTER(hcg_path_result)), _dp(sol), _dp(last), _dp(val),
                                logp))
    log = None
    if log_arr is not None:
        log = log_arr[: min(keep[-2].value, log_capacity)]
    return BatchOutput(rec, sol, last, val, log)


def _kernel_info(h):
    """VGPRs and private-segment bytes per lane, LDS bytes, waves per SIMD built for,
    workgroup size and CUs of a program's tracker kernel (hcg_program_kernel_info)."""
    out = (ctypes.c_int32 * 6)()
    check(lib().hcg_program_kernel_info(h, out))
    return dict(vgprs=int(out[0]), private_bytes_per_lane=int(out[1]), lds_bytes=int(out[2]),
                waves_per_simd_built=int(out[3]), workgroup=int(out[4]), cus=int(out[5]))


class MultiProgram:
    """Several homotopies compiled into one program (hcg_program_create_multi): one launch
    tracks the paths of all of them from a single work queue."""

    def __init__(self, ctx: GpuContext, homotopies, waves_per_eu: int = 0, layout: str = "lane"):
        if not 1 <= len(homotopies) <= 8:
            raise ValueError("between 1 and 8 homotopies per program")
        self.ctx = ctx
        self.homotopies = homotopies
        self.ns = [h.n for h in homotopies]
        self.waves_per_eu = waves_per_eu
        self.layout = layout
        self._h = _create(ctx, [h.desc() for h in homotopies], True, waves_per_eu, layout)

    def kernel_info(self):
        return _kernel_info(self._h)

    def close(self):
        if self._h:
            lib().hcg_program_destroy(self._h)
            self._h = ctypes.c_void_p()


def track_multi(program: MultiProgram, starts_list, opts: Optional[hcg_options] = None, target_groups_list=None,
                group_of_path_list=None):
    """Track the start solutions of every system of `program` in one launch; returns one
    BatchOutput per system."""
    opts = opts or abi.default_options()
    S = len(program.homotopies)
    if len(starts_list) != S:
        raise ValueError("one array of start solutions per system")
    batches = (hcg_batch * S)()
    keep = []
    total = totx = 0
    for s, (H, st) in enumerate(zip(program.homotopies, starts_list)):
        st = np.ascontiguousarray(np.asarray(st, dtype=np.complex128))
        if st.ndim != 2 or st.shape[1] != H.n:
            raise ValueError(f"system {s}: start solutions must be npaths x {H.n}")
        keep.append(st)
        b = batches[s]
        b.npaths = st.shape[0]
        b.starts = _dp(st)
        b.t1, b.t0 = 1.0, 0.0
        if getattr(H, "q", None) is not None:
            tg = None if target_groups_list is None else target_groups_list[s]
            p = np.ascontiguousarray(H.p)
            q = np.ascontiguousarray(H.q if tg is None else np.asarray(tg, dtype=np.complex128))
            keep += [p, q]
            b.start_params, b.target_params = _dp(p), _dp(q)
            b.ngroups = q.shape[0]
            gp = None if group_of_path_list is None else group_of_path_list[s]
            if gp is not None:
                g = np.ascontiguousarray(np.asarray(gp, dtype=np.int32))
                if g.shape != (st.shape[0],):
                    raise ValueError("group_of_path must have one entry per path")
                keep.append(g)
                b.group_of_path = g.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        total += st.shape[0]
        totx += st.shape[0] * H.n
    rec = np.zeros(total, dtype=_REC_DTYPE)
    sol = np.zeros(totx, dtype=np.complex128)
    last = np.zeros(totx, dtype=np.complex128)
    val = np.zeros(totx, dtype=np.float64)
    L = lib()
    L.hcg_track_multi.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(hcg_batch),
                                  ctypes.POINTER(hcg_options), ctypes.POINTER(hcg_path_result),
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)]
    check(L.hcg_track_multi(program.ctx.handle, program._h, batches, ctypes.byref(opts),
                            rec.ctypes.data_as(ctypes.POINTER(hcg_path_result)), _dp(sol), _dp(last), _dp(val)))
    outs = []
    off = offx = 0
    for s, H in enumerate(program.homotopies):
        np_ = batches[s].npaths
        n = H.n
        outs.append(BatchOutput(rec[off:off + np_], sol[offx:offx + np_ * n].reshape(np_, n),
                                last[offx:offx + np_ * n].reshape(np_, n), val[offx:offx + np_ * n].reshape(np_, n)))
        off += np_
        offx += np_ * n
    return outs


_CERT_DTYPE = np.dtype([(f, np.int32 if t is ctypes.c_int32 else np.float64) for f, t in abi.hcg_certificate._fields_],
                       align=True)
assert _CERT_DTYPE.itemsize == ctypes.sizeof(abi.hcg_certificate)


class CertifyOutput:
    """Per-candidate arrays of one hcg_certify call: the hcg_certificate fields, the boxes
    x₁ (box1) and x₀ (box0) as (nsol, n, 4) arrays of (re lo, re hi, im lo, im hi), and the
    refined candidates."""

    def __init__(self, rec, box1, box0, refined):
        self.records = rec
        for f in rec.dtype.names:
            setattr(self, f, rec[f])
        self.box1, self.box0, self.refined = box1, box0, refined


class Certifier:
    """hcg_certifier_create: a square system's point, double-double and interval
    evaluations plus the Krawczyk kernel, compiled with hipRTC."""

    def __init__(self, ctx: GpuContext, system, fixed_params=None):
        from .homotopies import _SysTapes
        self.ctx = ctx
        self.system = system
        self.n = system.nvars
        if len(system) != self.n:
            raise ValueError("only square systems can be certified")
        self.nparams = 0 if fixed_params is not None else system.nparams
        self._tapes = _SysTapes(system, fixed_params if system.nparams else None)
        self._struct = self._tapes.struct()
        self._h = ctypes.c_void_p()
        check(lib().hcg_certifier_create(ctx.handle, ctypes.byref(self._struct), self.n, ctypes.byref(self._h)))

    def certify(self, candidates, params=None, params_ival=None, params_ball=None, refine=True, extended=True,
                real_system=True, group_of=None) -> CertifyOutput:
        """params (and params_ival, params_ball): one parameter group (nparams, ...) or several
        (ngroups, nparams, ...); group_of: the group of each candidate (default group 0)."""
        X = np.ascontiguousarray(np.asarray(candidates, dtype=np.complex128).reshape(-1, self.n))
        nsol = len(X)
        rec = np.zeros(nsol, dtype=_CERT_DTYPE)
        box1 = np.zeros((nsol, self.n, 4))
        box0 = np.zeros((nsol, self.n, 4))
        ref = np.zeros((nsol, self.n), dtype=np.complex128)
        P = IP = BP = G = None
        ngroups = 0
        if self.nparams:
            if params is None:
                raise ValueError("the system has parameters: pass params")
            P = np.ascontiguousarray(np.asarray(params, dtype=np.complex128).reshape(-1, self.nparams))
            ngroups = len(P)
            if params_ival is not None:
                IP = np.ascontiguousarray(np.asarray(params_ival, dtype=np.float64).reshape(ngroups, self.nparams, 4))
            if params_ball is not None:
                BP = np.ascontiguousarray(np.asarray(params_ball, dtype=np.float64).reshape(ngroups, self.nparams, 6))
            if group_of is not None:
                G = np.ascontiguousarray(np.asarray(group_of, dtype=np.int32).reshape(nsol))
        keep = (X, P, IP, BP, G)  # noqa: F841  (alive during the call)
        opt = lambda a: _dp(a) if a is not None else None  # noqa: E731
        check(lib().hcg_certify(self.ctx.handle, self._h, nsol, _dp(X), ngroups, opt(P), opt(IP), opt(BP),
                                G.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if G is not None else None,
                                int(refine), int(extended), int(real_system),
                                rec.ctypes.data_as(ctypes.POINTER(abi.hcg_certificate)), _dp(box1), _dp(box0),
                                _dp(ref)))
        return CertifyOutput(rec, box1, box0, ref)

    def source(self) -> str:
        return lib().hcg_program_source(self._h).decode()

    def close(self):
        if self._h:
            lib().hcg_program_destroy(self._h)
            self._h = ctypes.c_void_p()
