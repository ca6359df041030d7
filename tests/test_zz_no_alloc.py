"""hcg_track_batch_device_noalloc (launch state in the library's fixed device scratch, no
allocation) returns exactly what hcg_track_batch_device returns for the same device-resident
batch: every record, endpoint, last point and valuation bit for bit and the same launch
counters, on two calls in a row, the second with a smaller batch (the scratch keeps the
first call's contents and must be cleared), with the plain and the learned proposer."""
import ctypes
from pathlib import Path

import numpy as np
import pytest
import torch

from hcgpu import abi, learned
from hcgpu.abi import hcg_batch, hcg_path_result
from hcgpu.configs import cfg2_problem
from hcgpu.homotopies import ParameterHomotopy
from hcgpu.solve import make_options

pytestmark = pytest.mark.gpu
STARTS = Path(__file__).resolve().parent / "golden" / "cfg2_starts.npy"


class _Dev:
    def __init__(self, H, starts, groups):
        dev = torch.device("cuda:0")
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        self.n = H.n
        self.npaths = len(starts)
        self.starts = t(np.ascontiguousarray(starts).view(np.float64))
        self.p = t(np.ascontiguousarray(H.p).view(np.float64))
        self.q = t(np.ascontiguousarray(H.q).view(np.float64))
        self.g = t(groups.astype(np.int32))
        torch.cuda.synchronize()
        b = hcg_batch()
        P = ctypes.POINTER(ctypes.c_double)
        b.npaths = self.npaths
        b.starts = ctypes.cast(self.starts.data_ptr(), P)
        b.start_params = ctypes.cast(self.p.data_ptr(), P)
        b.target_params = ctypes.cast(self.q.data_ptr(), P)
        b.ngroups = H.q.shape[0]
        b.group_of_path = ctypes.cast(self.g.data_ptr(), ctypes.POINTER(ctypes.c_int32))
        b.t1, b.t0 = 1.0, 0.0
        self.b = b

    def run(self, ctx, prog, opts, entry):
        dev = torch.device("cuda:0")
        res = torch.full((self.npaths * ctypes.sizeof(hcg_path_result),), 0xAB, dtype=torch.uint8, device=dev)
        sol = torch.full((self.npaths * self.n * 2,), 7.0, dtype=torch.float64, device=dev)
        last = torch.full_like(sol, 7.0)
        val = torch.full((self.npaths * self.n,), 7.0, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        L = abi.lib()
        fn = getattr(L, entry)
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(hcg_batch), ctypes.POINTER(abi.hcg_options),
                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        abi.check(fn(ctx.handle, prog._h, ctypes.byref(self.b), ctypes.byref(opts), res.data_ptr(), sol.data_ptr(),
                     last.data_ptr(), val.data_ptr()))
        torch.cuda.synchronize()
        return (res.cpu().numpy(), sol.cpu().numpy(), last.cpu().numpy(), val.cpu().numpy(), ctx.last_counters())


@pytest.mark.parametrize("use_learned", [0, 1])
def test_noalloc_equals_device_entry(gpu_ctx, use_learned):
    F, p0, Q = cfg2_problem(4)
    S = np.load(STARTS)
    H = ParameterHomotopy(F, p0, Q, lift_affine=True)
    prog = gpu_ctx.program(H)
    if use_learned:
        learned.load(gpu_ctx, learned.saved_weights("param10_step_mlp"))
    opts = make_options(learned_step=use_learned)
    for ntargets in (4, 1):  # the second call is smaller than the first
        starts = np.tile(S, (ntargets, 1))
        groups = np.repeat(np.arange(ntargets, dtype=np.int32), len(S))
        d = _Dev(H, starts, groups)
        ref = d.run(gpu_ctx, prog, opts, "hcg_track_batch_device")
        got = d.run(gpu_ctx, prog, opts, "hcg_track_batch_device_noalloc")
        print(f"{len(starts)} paths: counters {got[4]}")
        for a, b in zip(ref[:4], got[:4]):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        assert ref[4] == got[4]
        rec = np.frombuffer(got[0].tobytes(), dtype=np.int32).reshape(len(starts), -1)
        assert (rec[:, 0] == 1).all()  # every path of this batch ends in success
