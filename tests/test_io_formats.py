"""Solution / parameter files in the reference's text format (src/utils.jl:3-134) and the
seed-directory layout of its path-tracker benchmark data.

Pinned by: the reference's docstring example (exact text), its round-trip test
(test/utils_test.jl:3-9, exact equality), and its own data file: four of the 3,264 real
conics of test/data/3264_real_sols.txt read back in coordinate order satisfy the Steiner
system at the five real conics of test/certification_test.jl:256-289.
"""
import json
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

import systems
from hcgpu import io

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_write_parameters_matches_reference_text(tmp_path):
    """The docstring example of write_parameters (src/utils.jl:90-104)."""
    f = tmp_path / "parameters.txt"
    io.write_parameters(f, [2.0, -3.2 + 2j])
    assert f.read_text() == "2\n\n2.0 0.0\n-3.2 2.0\n"
    np.testing.assert_array_equal(io.read_parameters(f), [2.0, -3.2 + 2j])


def test_write_solutions_layout(tmp_path):
    """The docstring example of write_solutions (src/utils.jl:17-27), float-valued."""
    f = tmp_path / "solutions.txt"
    io.write_solutions(f, [[1, 1], [-1, 2]])
    assert f.read_text() == "2\n\n1.0 0.0\n1.0 0.0\n\n-1.0 0.0\n2.0 0.0\n"
    np.testing.assert_array_equal(io.read_solutions(f), [[1, 1], [-1, 2]])


def test_round_trip_is_exact(tmp_path):
    """test/utils_test.jl:3-9: read(write(S)) == S and read(write(p)) == p, bit for bit."""
    rng = np.random.default_rng(3)
    S = rng.standard_normal((2, 3)) + 1j * rng.standard_normal((2, 3))
    S[0, 0] = 1e-300 - 2.5e300j
    io.write_solutions(tmp_path / "s.txt", S)
    R = io.read_solutions(tmp_path / "s.txt")
    assert R.tobytes() == S.tobytes()
    p = rng.standard_normal(14) + 1j * rng.standard_normal(14)
    io.write_parameters(tmp_path / "p.txt", p)
    assert io.read_parameters(tmp_path / "p.txt").tobytes() == p.tobytes()


def test_reference_data_file_satisfies_steiner_system():
    """Real tangent conics from the reference's data file solve the Steiner system at the
    five real conics of its certification test (coordinates in the file's order)."""
    sols = io.read_solutions(GOLDEN / "conics_3264_sample.txt")
    assert sols.shape == (4, 15)
    conics = [float(Fraction(c)) for c in json.loads((GOLDEN / "param_cases.json").read_text())["conics"]["real_conics"]]
    F = systems.steiner()[0]
    for s in sols:
        vals = list(s) + conics
        res = max(abs(f(vals)) for f in F.polys)
        scale = max(sum(abs(c) * abs(np.prod([v ** k for v, k in zip(vals, e)])) for e, c in f.terms.items())
                    for f in F.polys)
        assert res <= 1e-12 * scale, res


def test_malformed_files_are_errors(tmp_path):
    f = tmp_path / "bad.txt"
    f.write_text("3\n\n1.0 0.0\n\n2.0 0.0\n")
    with pytest.raises(ValueError):
        io.read_solutions(f)
    f.write_text("2\n\n1.0 0.0\n1.0 0.0\n\n2.0 0.0\n")
    with pytest.raises(ValueError):
        io.read_solutions(f)
    f.write_text("4\n\n1.0 0.0\n")
    with pytest.raises(ValueError):
        io.read_parameters(f)


def test_seed_directory_round_trip(tmp_path):
    """The three files of a seed directory (benchmarks/generate_path_tracker_seed_data.jl:59-72)
    describe the Steiner case: reading them back gives the batch that reaches s_q."""
    F, p, q, s_p, s_q = systems.steiner()
    io.write_seed_directory(tmp_path / "steiner" / "seed_1", p, [s_p], q)
    p2, S2, q2 = io.read_seed_directory(tmp_path / "steiner" / "seed_1")
    assert p2.tobytes() == p.astype(np.complex128).tobytes() and q2.tobytes() == q.astype(np.complex128).tobytes()
    assert S2.shape == (1, 15) and S2[0].tobytes() == s_p.tobytes()
