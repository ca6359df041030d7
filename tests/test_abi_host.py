"""C-ABI and host-side checks that need no GPU.

* libhcgpu.so loads and exports every function include/hcgpu.h declares;
* the ctypes mirrors of the ABI structs have the C compiler's layout (sizes and
  every field offset, from a probe compiled with gcc against include/hcgpu.h);
* the generator and hipRTC compile a program on the CPU, and report errors as
  status codes with a message (no exceptions across the ABI);
* there is no CPU fallback: without a GPU the context fails loudly, and nothing in
  the package imports the oracle;
* host logic: start-solution order, the tape producer, Result counting and
  multiplicity clustering.
"""
import ctypes
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from hcgpu import abi
from hcgpu.configs import dense_system
from hcgpu.homotopies import ParameterHomotopy, total_degree, total_degree_start_solutions
from hcgpu.polys import System, variables
from hcgpu.result import PathResult, Result, multiplicities
from hcgpu.tape import hcg_tape, tapes

REPO = Path(__file__).resolve().parent.parent
HEADER = REPO / "include" / "hcgpu.h"
PKG = REPO / "homotopycontinuation.jl_amd"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(hcg_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = abi.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(abi.HCGPU_SYMBOLS) <= set(names)


STRUCTS = {
    "hcg_tape": hcg_tape,
    "hcg_system": abi.hcg_system,
    "hcg_homotopy_desc": abi.hcg_homotopy_desc,
    "hcg_options": abi.hcg_options,
    "hcg_batch": abi.hcg_batch,
    "hcg_path_result": abi.hcg_path_result,
    "hcg_step_log": abi.hcg_step_log,
    "hcg_build_options": abi.hcg_build_options,
    "hcg_certificate": abi.hcg_certificate,
}


def test_ctypes_layout_matches_c_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "hcgpu.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'printf("{cname}.{f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", str(REPO / "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                              check=True).stdout.splitlines())
    for cname, cls in STRUCTS.items():
        assert int(out[cname]) == ctypes.sizeof(cls), cname
        for f in cls._fields_:
            assert int(out[f"{cname}.{f[0]}"]) == getattr(cls, f[0]).offset, f"{cname}.{f[0]}"


def _small_homotopy():
    (x, y), _ = variables(2)
    F = System([2.3 * x ** 2 + 1.2 * y ** 2 + 3 * x - 2 * y + 3, 2.3 * x ** 2 + 1.2 * y ** 2 + 5 * x + 2 * y - 5], 2)
    return total_degree(F, gamma=1.3j + 0.4)


def test_generate_and_hiprtc_compile_on_cpu():
    L = abi.lib()
    H, _ = _small_homotopy()
    need = ctypes.c_size_t()
    fl = (ctypes.c_long * 6)()
    rc = L.hcg_generate(ctypes.byref(H.desc()), None, 0, ctypes.byref(need), fl)
    assert rc == 0 and need.value > 1000
    buf = ctypes.create_string_buffer(need.value)
    assert L.hcg_generate(ctypes.byref(H.desc()), buf, need.value, ctypes.byref(need), fl) == 0
    src = buf.value.decode()
    for macro in ["HCG_N 2", "HCG_KIND", "HCG_FL_EJ", "HCG_FL_T3", "hcg_tracker.h"]:
        assert macro in src
    # eval, eval_jac, eval_dd, taylor 1..3 of the system functions (order 1 of a straight-line
    # homotopy is γG - F, made from eval, so its own Taylor count stays 0)
    assert fl[0] > 0 and fl[1] > fl[0] and fl[2] > fl[0] and fl[5] > fl[4] > 0, list(fl)
    size = ctypes.c_size_t()
    assert L.hcg_compile_only(ctypes.byref(H.desc()), ctypes.byref(size)) == 0, L.hcg_last_error()
    assert size.value > 10000  # a gfx950 code object


def test_invalid_tape_is_an_error_code_with_message():
    L = abi.lib()
    H, _ = _small_homotopy()
    d = H.desc()
    t = d.target.eval_jac.contents
    saved = t.ninstr
    t.ninstr = -3
    try:
        size = ctypes.c_size_t()
        rc = L.hcg_compile_only(ctypes.byref(d), ctypes.byref(size))
        assert rc != 0
        assert L.hcg_last_error()
    finally:
        t.ninstr = saved


def test_parameter_homotopy_requires_parameters():
    L = abi.lib()
    F = dense_system(2, 2, seed=1)
    H, _ = total_degree(F, seed=1)
    d = abi.hcg_homotopy_desc()
    d.kind = abi.HCG_PARAMETER
    d.n = 2
    d.target = H.F.struct()
    size = ctypes.c_size_t()
    assert L.hcg_compile_only(ctypes.byref(d), ctypes.byref(size)) != 0
    assert b"parameter" in L.hcg_last_error()


def test_graph_entry_points_reject_missing_arguments():
    """hcg_graph_create / hcg_graph_launch validate before touching a device; destroy of NULL is a no-op."""
    L = abi.lib()
    out = ctypes.c_void_p()
    rc = L.hcg_graph_create(None, None, None, None, None, None, None, None, ctypes.byref(out))
    assert rc == abi.HCG_ERR_INVALID_ARGUMENT and not out.value
    assert L.hcg_graph_create(None, None, None, None, None, None, None, None, None) == abi.HCG_ERR_INVALID_ARGUMENT
    assert L.hcg_graph_launch(None) == abi.HCG_ERR_INVALID_ARGUMENT
    assert b"graph" in L.hcg_last_error()
    L.hcg_graph_destroy(None)


def test_loading_the_library_does_not_import_torch():
    """The code-object cache is compiled in processes that load libhcgpu.so without torch, so
    that /opt/rocm's hipRTC compiles it (torch bundles an older hipRTC under the same soname,
    whose code ran the parameter workload 27% slower)."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import hcgpu; from hcgpu import abi; abi.lib(); "
            "print('torch' in sys.modules)" % str(REPO / "homotopycontinuation.jl_amd"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "False"


def test_no_gpu_means_a_loud_failure():
    """The product path has no CPU fallback: without a device the context cannot be made."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from hcgpu.gpu import GpuContext
    with pytest.raises(abi.HcgpuError):
        GpuContext(0)


def test_package_never_imports_the_oracle():
    for f in PKG.rglob("*.py"):
        text = f.read_text()
        assert not re.search(r"^\s*(import|from)\s+oracle\b", text, flags=re.M), f
    for f in list(PKG.rglob("*.cpp")) + list(PKG.rglob("*.h")) + list(PKG.rglob("*.hip")):
        assert "hc_oracle" not in f.read_text(), f


# ---------------------------------------------------------------- host logic
def test_total_degree_start_solutions_order():
    """Iterators.product order: the first coordinate varies fastest (src/total_degree.jl:235-262)."""
    S = total_degree_start_solutions([2, 3])
    assert S.shape == (6, 2)
    np.testing.assert_allclose(S[0], [1, 1], atol=1e-15)
    np.testing.assert_allclose(S[1], [-1, 1], atol=1e-15)
    np.testing.assert_allclose(S[2], [1, np.exp(2j * np.pi / 3)], atol=1e-15)
    S = total_degree_start_solutions([3, 3, 2])
    assert len({tuple(np.round(s, 12)) for s in S}) == 18
    np.testing.assert_allclose(S[:, 0] ** 3, 1, atol=1e-14)
    np.testing.assert_allclose(S[:, 2] ** 2, 1, atol=1e-14)


def test_total_degree_start_system_scaling():
    """G_i = s_i (x_i^{d_i} - 1), s_i = max_α |c_iα| (src/total_degree.jl:62, 102-106)."""
    F = dense_system(3, 3, seed=4)
    H, S = total_degree(F, seed=5)
    s = [max(abs(c) for c in f.terms.values()) for f in F.polys]
    np.testing.assert_allclose(H.G.fixed.view(np.complex128).real, s)
    assert abs(abs(H.gamma) - 1) < 1e-15


def test_tape_struct_fields():
    """The tape mirrors InstructionSequence (src/model_kit/instruction_sequence.jl:1-27): 6 int32
    per instruction, 1-based slots, constants / parameters / variables blocks in that order."""
    (x, y), (a,) = variables(2, 1)
    F = System([x * y - a, x ** 2 + 3.5], 2, 1)
    ej, ev = tapes(F)
    t = ej.struct()
    assert t.nvars == 2 and t.nparams == 1
    assert t.params_first == t.nconst + 1 and t.vars_first == t.params_first + 1
    instr = np.ctypeslib.as_array(t.instr, shape=(t.ninstr, 6))
    assert (instr[:, 5] >= t.vars_first + t.nvars).all()
    assert t.nu == 2 and t.nU >= 3  # ∂(xy - a) = (y, x), ∂(x² + 3.5) = (2x, 0)
    assert ev.struct().nU == 0


def test_parameter_homotopy_groups():
    (x,), (a,) = variables(1, 1)
    F = System([x ** 2 - a], 1, 1)
    H = ParameterHomotopy(F, [1.0], [[2.0], [3.0], [4.0]])
    assert H.q.shape == (3, 1)
    with pytest.raises(ValueError):
        ParameterHomotopy(F, [1.0, 2.0], [2.0])


def _pr(code, x, singular=False, winding=0):
    return PathResult(return_code=code, solution=np.asarray(x, dtype=complex), singular=singular,
                      winding_number=winding or None, accuracy=1e-15, residual=1e-15, condition_jacobian=1.0,
                      t=0.0, accepted_steps=10, rejected_steps=0, extended_precision=False, path_number=0,
                      start_solution=None, last_path_point=None, valuation=None, ω=0.0, μ=0.0,
                      extended_precision_used=False, steps_eg=0, newton_iters=0)


def test_multiplicity_clustering():
    """src/unique_points.jl:314-340 semantics: atol 1e-14 or rtol 1e-8 (Euclidean)."""
    pts = [np.array([1.0, 2.0]), np.array([1.0 + 1e-9, 2.0]), np.array([5.0, 5.0]), np.array([1.0, 2.0 + 5e-9])]
    groups = multiplicities(pts)
    assert sorted(map(sorted, groups)) == [[0, 1, 3]]


def test_result_counts_singular_clusters_once():
    """nsingular counts clusters once; at_infinity counts separately (src/result.jl:627-841)."""
    prs = [_pr("success", [1.0, 2.0]), _pr("success", [3.0, 1.0]),
           _pr("success", [1.0, 1.0], singular=True, winding=2), _pr("success", [1.0, 1.0 + 1e-10], singular=True, winding=2),
           _pr("at_infinity", [1e9, 1e9]), _pr("terminated_max_steps", [0.0, 0.0])]
    for k, r in enumerate(prs):
        r.path_number = k + 1
    R = Result(prs)
    assert R.ntracked() == 6
    assert R.nnonsingular() == 2
    assert R.nsingular() == 1
    assert R.nsingular(counting_multiplicities=True) == 2
    assert R.nat_infinity() == 1
    assert R.nfailed() == 1
    assert R.nresults() == 3


def test_certifier_compile_errors_are_status_codes():
    """hcg_certifier_compile_only reports bad systems as error codes with a message."""
    import ctypes
    from hcgpu.homotopies import _SysTapes
    (x, y), _ = variables(2, 0)
    F = System([x**2 + y**2 - 1, x - y], 2, 0)
    tapes = _SysTapes(F)
    st = tapes.struct()
    L = abi.lib()
    size = ctypes.c_size_t()
    assert L.hcg_certifier_compile_only(ctypes.byref(st), 3, ctypes.byref(size)) != 0  # wrong size
    assert b"" != L.hcg_last_error()
    assert L.hcg_certifier_compile_only(None, 2, ctypes.byref(size)) != 0
    assert L.hcg_certifier_compile_only(ctypes.byref(st), 2, ctypes.byref(size)) == 0 and size.value > 0
