"""Writes param_cases.json: the explicit parameter and solution vectors of the reference's
parameter-homotopy test cases, transcribed from its test files:

  steiner   test/test_cases/steiner_higher_prec.jl:5-105  (p, q, s_p, s_q; asserts :141-152)
  four_bar  test/test_cases/four_bar.jl:27-88             (s, p, q; asserts :90-97)
  conics    test/certification_test.jl:256-289: the 30 rational coefficients of five real
            conics (exact fractions), and conics_3264_sample.txt: the first 4 of the 3,264
            real tangent conics of test/data/3264_real_sols.txt, in the same file format with
            the count line set to 4; conics_3264_real.npy: all 3,264 of them (3264 x 15
            complex128, the file's values), certified in test/certification_test.jl:290-295.

The vectors are data (inputs and expected outputs); the systems themselves are rebuilt in
tests/systems.py.  Run from the repository root where the reference checkout exists:

    python tests/golden/extract_param_cases.py /path/to/reference
"""
import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
NUM = r"[-+]?\d+\.\d+(?:e[-+]?\d+)?"
UNUM = r"\d+\.\d+(?:e[-+]?\d+)?"


def vector(text: str, name: str):
    """Entries of `name = [ ... ]` (optionally `Complex{Float64}[ ... ]`), as [re, im] pairs."""
    m = re.search(r"\b" + re.escape(name) + r"\s*=\s*(?:Complex\{Float64\})?\[(.*?)\]", text, re.S)
    if not m:
        raise ValueError(f"{name} not found")
    out = []
    for item in m.group(1).split(","):
        item = item.strip()
        if not item:
            continue
        c = re.fullmatch(rf"({NUM})\s*([-+])\s*({UNUM})im", item)
        if c:
            im = float(c.group(3)) * (-1 if c.group(2) == "-" else 1)
            out.append([float(c.group(1)), im])
        else:
            out.append([float(item), 0.0])
    return out


def main(ref: Path):
    st = (ref / "test" / "test_cases" / "steiner_higher_prec.jl").read_text()
    fb = (ref / "test" / "test_cases" / "four_bar.jl").read_text()
    data = {
        "steiner": {k: vector(st, k) for k in ("p", "q", "s_p", "s_q")},
        "four_bar": {k: vector(fb, k) for k in ("s", "p", "q")},
    }
    cert = (ref / "test" / "certification_test.jl").read_text()
    block = re.search(r"real_conics\s*=\s*\[(.*?)\]", cert, re.S).group(1)
    data["conics"] = {"real_conics": re.findall(r"-?\d+\s*//\s*\d+", block)}
    data["conics"]["real_conics"] = [c.replace(" ", "").replace("//", "/") for c in data["conics"]["real_conics"]]
    assert len(data["conics"]["real_conics"]) == 30
    lines = (ref / "test" / "data" / "3264_real_sols.txt").read_text().splitlines()
    nsol = 4
    (HERE / "conics_3264_sample.txt").write_text("\n".join([str(nsol)] + lines[1:1 + 16 * nsol]) + "\n")
    sys.path.insert(0, str(HERE.parents[1] / "homotopycontinuation.jl_amd"))
    from hcgpu.io import read_solutions
    import numpy as np
    allsols = read_solutions(ref / "test" / "data" / "3264_real_sols.txt")
    assert allsols.shape == (3264, 15)
    np.save(HERE / "conics_3264_real.npy", allsols)
    assert [len(data["steiner"][k]) for k in ("p", "q", "s_p", "s_q")] == [30, 30, 15, 15]
    assert [len(data["four_bar"][k]) for k in ("s", "p", "q")] == [24, 16, 16]
    (HERE / "param_cases.json").write_text(json.dumps(data, indent=1))


if __name__ == "__main__":
    main(Path(sys.argv[1]))
