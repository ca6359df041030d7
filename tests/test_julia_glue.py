"""The Julia side of the boundary (julia/HCGPU.jl) against the C headers it binds.

No Julia toolchain is available, so the mirror is checked textually: every C struct of
include/hcgpu.h and include/hcgpu_types.h has a Julia struct with the same fields in the same
order and with the corresponding types (so Julia's C-compatible isbits layout equals the C
layout); every entry point has exactly one ccall with the same name, return type, arity and
argument types; every numeric #define has a Julia constant of the same value.
"""
import re
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
HEADERS = [REPO / "include" / "hcgpu.h", REPO / "include" / "hcgpu_types.h"]
JULIA = REPO / "julia" / "HCGPU.jl"

SCALARS = {"int32_t": "Int32", "double": "Float64", "float": "Float32", "uint64_t": "UInt64",
           "int64_t": "Int64", "size_t": "Csize_t", "long": "Clong", "int": "Cint", "char": "UInt8",
           "void": "Cvoid"}
OPAQUE = {"hcg_ctx", "hcg_program", "hcg_graph"}


def strip_c_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def camel(c_name):
    return "".join(w.capitalize() for w in c_name.split("_"))


def julia_type(base, nptr, array=None):
    if base in OPAQUE:
        assert nptr >= 1, base
        t = "Cvoid"
    elif base in SCALARS:
        t = SCALARS[base]
    else:
        assert base.startswith("hcg_"), base
        t = camel(base)
    for _ in range(nptr):
        t = f"Ptr{{{t}}}"
    if array is not None:
        t = f"NTuple{{{array},{t}}}"
    return t


def c_header_text():
    return "\n".join(strip_c_comments(p.read_text()) for p in HEADERS)


def c_structs():
    out = {}
    for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} (\w+);", c_header_text(), re.S):
        fields = []
        for decl in m.group(2).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            dm = re.match(r"^(?:const\s+)?(\w+)\s*(\**)\s*(.+)$", decl, re.S)
            assert dm, decl
            base, star = dm.group(1), len(dm.group(2))
            for d in dm.group(3).split(","):
                vm = re.match(r"^\s*(\**)\s*(\w+)\s*(?:\[(\d+)\])?\s*$", d)
                assert vm, d
                fields.append((vm.group(2), julia_type(base, star + len(vm.group(1)), vm.group(3))))
        out[m.group(1)] = fields
    return out


def c_functions():
    out = {}
    text = c_header_text()
    for m in re.finditer(r"(?:^|[;}\n])\s*((?:const\s+)?\w+\s*\**)\s*\b(hcg_\w+)\s*\(([^)]*)\)\s*;", text, re.S):
        ret, name, params = m.group(1), m.group(2), m.group(3).strip()
        rm = re.match(r"^(const\s+)?(\w+)\s*(\**)$", ret.strip())
        rbase, rptr = rm.group(2), len(rm.group(3))
        if rbase == "char" and rptr == 1:
            rtype = "Cstring"
        else:
            rtype = julia_type(rbase, rptr)
        args = []
        if params and params != "void":
            for p in params.split(","):
                p = " ".join(p.split())
                nptr = p.count("*")
                words = [w for w in re.split(r"[\s\*]+", p) if w and w != "const"]
                base = words[0]
                args.append(julia_type(base, nptr))
        out[name] = (rtype, args)
    return out


def c_defines():
    out = {}
    for p in HEADERS:
        for m in re.finditer(r"^#define (HCG_\w+) \(?(-?\d+)\)?", p.read_text(), re.M):
            out[m.group(1)] = int(m.group(2))
    return out


def julia_text():
    return JULIA.read_text()


def julia_structs():
    out = {}
    for m in re.finditer(r"^(?:mutable\s+)?struct\s+(Hcg\w+)\s*$(.*?)^end\b", julia_text(), re.M | re.S):
        fields = []
        for line in m.group(2).splitlines():
            line = line.split("#", 1)[0].strip()
            for part in filter(None, (s.strip() for s in line.split(";"))):
                fm = re.match(r"^(\w+)::(.+)$", part)
                assert fm, part
                fields.append((fm.group(1), fm.group(2).replace(" ", "")))
        out[m.group(1)] = fields
    return out


def julia_ccalls():
    out = {}
    for m in re.finditer(r"ccall\(\(:(\w+),\s*libhcgpu\),\s*(\w+),\s*\((.*?)\)\s*,?", julia_text(), re.S):
        types = [t.strip() for t in m.group(3).split(",") if t.strip()]
        assert m.group(1) not in out, f"two ccalls of {m.group(1)}"
        out[m.group(1)] = (m.group(2), types)
    return out


def test_headers_parse():
    s = c_structs()
    assert {"hcg_tape", "hcg_system", "hcg_homotopy_desc", "hcg_batch", "hcg_options", "hcg_path_result",
            "hcg_step_log", "hcg_certificate", "hcg_build_options"} <= set(s)
    assert s["hcg_tape"][0] == ("instr", "Ptr{Int32}")
    assert s["hcg_build_options"][-1] == ("reserved", "NTuple{6,Int32}")
    f = c_functions()
    assert f["hcg_ctx_create"] == ("Cint", ["Cint", "Ptr{Ptr{Cvoid}}"])
    assert f["hcg_program_create_multi"][1][1] == "Ptr{Ptr{HcgHomotopyDesc}}"
    assert f["hcg_last_error"] == ("Cstring", [])
    # every hcg_* name followed by an argument list is a parsed declaration
    assert set(f) == set(re.findall(r"\b(hcg_\w+)\s*\(", c_header_text()))


@pytest.mark.parametrize("c_name", sorted(c_structs()))
def test_struct_mirror(c_name):
    js = julia_structs()
    jname = camel(c_name)
    assert jname in js, f"julia/HCGPU.jl has no struct {jname} mirroring {c_name}"
    assert js[jname] == c_structs()[c_name], f"{jname} fields differ from {c_name}"


def test_no_extra_julia_structs():
    assert set(julia_structs()) == {camel(c) for c in c_structs()}


def test_every_entry_point_has_one_ccall_with_the_c_signature():
    cf, jc = c_functions(), julia_ccalls()
    assert set(cf) == set(jc), (sorted(set(cf) - set(jc)), sorted(set(jc) - set(cf)))
    for name, (ret, args) in cf.items():
        jret, jargs = jc[name]
        assert jret == ret, (name, jret, ret)
        assert jargs == args, (name, jargs, args)


def test_library_exports_what_the_julia_side_binds():
    import sys
    sys.path.insert(0, str(REPO / "homotopycontinuation.jl_amd"))
    from hcgpu import abi
    missing = set(julia_ccalls()) - set(abi.HCGPU_SYMBOLS)
    assert not missing, missing


def test_constants():
    text = julia_text()
    consts = {m.group(1): m.group(2) for m in re.finditer(r"^const (HCG_\w+) = (.+?)$", text, re.M)}
    for name, value in c_defines().items():
        assert name in consts, name
        assert int(consts[name]) == value, (name, consts[name], value)
    # the one computed constant: the parameter count of the step network
    expr = re.search(r"^const HCG_MLP_NPARAMS = (.*?)\n\n", text, re.M | re.S).group(1)
    env = {k: int(v) for k, v in consts.items() if re.fullmatch(r"-?\d+", v)}
    assert eval(expr.replace("\n", " "), {}, env) == 4866


def test_return_codes_follow_the_c_ordinals():
    # the C records carry EndgameTrackerCode ordinals; the Julia side maps them by position
    text = julia_text()
    assert "const CODES = instances(HC.EndgameTrackerCode.codes)" in text
    assert "Symbol(CODES[r.code+1])" in text
