"""Device-resident batches: inputs and outputs held in HBM (torch allocations on the GPU),
tracked with hcg_track_batch_device, or captured once as a HIP graph (hcg_graph_create) and
replayed batch after batch through the same buffers (hcg_graph_launch)."""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
# torch before libhcgpu.so: torch bundles its own HIP runtime under the sonames of /opt/rocm's;
# imported first, it is the runtime the library binds to as well.  The other way round the
# process holds two runtimes and torch finds no GPU.
import torch

from . import abi
from .abi import check, hcg_batch, hcg_options, hcg_path_result

_REC_DTYPE = np.dtype([(f, np.int32 if t is ctypes.c_int32 else np.float64) for f, t in hcg_path_result._fields_],
                      align=True)


class DeviceBatch:
    """A batch whose starts, parameters, group indices and outputs live in device memory."""

    def __init__(self, ctx, prog, starts, p=None, Q=None, groups=None, device: int = 0):
        self._torch = torch
        dev = torch.device(f"cuda:{device}")
        try:
            torch.cuda.init()
        except RuntimeError as e:
            raise abi.HcgpuError("torch cannot reach the GPU: import torch (or hcgpu.device) before the first "
                                 "GpuContext, so that torch and libhcgpu.so share one HIP runtime") from e
        starts = np.ascontiguousarray(starts, dtype=np.complex128)
        self.npaths, self.n = starts.shape
        self.ctx, self.prog = ctx, prog

        def t(a):
            return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(dev)

        self.starts = t(starts.view(np.float64))
        self.p = t(None if p is None else np.ascontiguousarray(p, dtype=np.complex128).view(np.float64))
        self.q = t(None if Q is None else np.ascontiguousarray(Q, dtype=np.complex128).view(np.float64))
        if Q is not None and groups is None:
            groups = np.zeros(self.npaths, dtype=np.int32)
        self.g = t(None if groups is None else np.asarray(groups, dtype=np.int32))
        self.res = torch.empty(self.npaths * ctypes.sizeof(hcg_path_result), dtype=torch.uint8, device=dev)
        self.sol = torch.empty(self.npaths * self.n * 2, dtype=torch.float64, device=dev)
        self.last = torch.empty_like(self.sol)
        self.val = torch.empty(self.npaths * self.n, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        b = hcg_batch()
        b.npaths = self.npaths
        P = ctypes.POINTER(ctypes.c_double)
        b.starts = ctypes.cast(self.starts.data_ptr(), P)
        b.t1, b.t0 = 1.0, 0.0
        if self.p is not None:
            b.start_params = ctypes.cast(self.p.data_ptr(), P)
            b.target_params = ctypes.cast(self.q.data_ptr(), P)
            b.ngroups = Q.shape[0]
            b.group_of_path = ctypes.cast(self.g.data_ptr(), ctypes.POINTER(ctypes.c_int32))
        self.b = b

    def _outputs(self):
        return self.res.data_ptr(), self.sol.data_ptr(), self.last.data_ptr(), self.val.data_ptr()

    def launch(self, opts: hcg_options):
        """One hcg_track_batch_device launch; returns (kernel ms, launch counters)."""
        check(abi.lib().hcg_track_batch_device(self.ctx.handle, self.prog._h, ctypes.byref(self.b),
                                               ctypes.byref(opts), *self._outputs()))
        return self.ctx.last_kernel_ms(), self.ctx.last_counters()

    def graph(self, opts: hcg_options) -> "TrackGraph":
        """The same launch captured as a HIP graph over these buffers."""
        return TrackGraph(self, opts)

    def set_starts(self, starts):
        """Overwrite the start solutions in place (same path count)."""
        a = np.ascontiguousarray(starts, dtype=np.complex128)
        if a.shape != (self.npaths, self.n):
            raise ValueError(f"expected starts of shape {(self.npaths, self.n)}, got {a.shape}")
        self.starts.copy_(self._torch.from_numpy(a.view(np.float64)).reshape(self.starts.shape))

    def set_targets(self, Q):
        """Overwrite the target parameters in place (same group count)."""
        a = np.ascontiguousarray(Q, dtype=np.complex128).view(np.float64).reshape(-1)
        if self.q is None or a.size != self.q.numel():
            raise ValueError("target parameters of another shape")
        self.q.copy_(self._torch.from_numpy(a).reshape(self.q.shape))

    def results(self):
        """(records, solutions, last points, valuations) copied to the host."""
        rec = np.frombuffer(self.res.cpu().numpy().tobytes(), dtype=_REC_DTYPE)
        sol = self.sol.cpu().numpy().view(np.complex128).reshape(self.npaths, self.n)
        last = self.last.cpu().numpy().view(np.complex128).reshape(self.npaths, self.n)
        val = self.val.cpu().numpy().reshape(self.npaths, self.n)
        return rec, sol, last, val


class TrackGraph:
    """hcg_graph_create over a DeviceBatch's buffers; launch() replays it.  Close it (or let it
    be collected) before its program and context are destroyed."""

    def __init__(self, batch: DeviceBatch, opts: hcg_options):
        self.batch = batch
        self.opts = opts
        self._h: Optional[ctypes.c_void_p] = ctypes.c_void_p()
        check(abi.lib().hcg_graph_create(batch.ctx.handle, batch.prog._h, ctypes.byref(batch.b), ctypes.byref(opts),
                                         *batch._outputs(), ctypes.byref(self._h)))

    def launch(self):
        """Replay the captured launch; returns (kernel ms, launch counters)."""
        if not self._h:
            raise abi.HcgpuError("graph is closed")
        check(abi.lib().hcg_graph_launch(self._h))
        return self.batch.ctx.last_kernel_ms(), self.batch.ctx.last_counters()

    def close(self):
        if self._h:
            abi.lib().hcg_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
