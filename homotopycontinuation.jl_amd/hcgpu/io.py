"""Solution and parameter files in the reference's plain-text (Bertini) format, and the
seed-directory layout its path-tracker benchmark data uses.

Formats (src/utils.jl:3-134):
  solutions   first line: the number of solutions; then, for each solution, a blank line
              followed by one "re im" line per coordinate (write_solutions / read_solutions).
  parameters  first line: the number of values, a blank line, then one "re im" line per
              value (write_parameters / read_parameters).
Numbers are written in the shortest form that reads back to the same double, as Julia's
`print` of a Float64 does, so a write followed by a read is exact.

Seed directory (benchmarks/generate_path_tracker_seed_data.jl:59-72): start_parameters.txt,
start_solutions.txt and target_parameters.txt describe one parameter-homotopy batch.
"""
from __future__ import annotations

from pathlib import Path
from typing import Iterable, List, Sequence, Tuple, Union

import numpy as np

PathLike = Union[str, Path]


def _num(v: float) -> str:
    """Shortest decimal string that parses back to the same double (as Julia prints it)."""
    v = float(v)
    if v == int(v) and abs(v) < 1e16:
        return f"{v:.1f}"
    return repr(v)


def _pair(z: complex) -> str:
    z = complex(z)
    return f"{_num(z.real)} {_num(z.imag)}"


def write_solutions(filename: PathLike, solutions: Iterable[Sequence[complex]]) -> None:
    """write_solutions(filename, S), src/utils.jl:35-43."""
    S = [np.asarray(s, dtype=np.complex128).ravel() for s in solutions]
    lines = [str(len(S))]
    for s in S:
        lines.append("")
        lines.extend(_pair(z) for z in s)
    Path(filename).write_text("\n".join(lines) + "\n")


def _rows(filename: PathLike) -> List[List[str]]:
    return [ln.split() for ln in Path(filename).read_text().splitlines()]


def read_solutions(filename: PathLike) -> np.ndarray:
    """read_solutions(filename), src/utils.jl:61-77: an (nsolutions, nvars) complex array."""
    rows = _rows(filename)
    if not rows or len(rows[0]) != 1:
        raise ValueError(f"{filename}: first line must hold the number of solutions")
    count = int(float(rows[0][0]))
    blocks: List[List[complex]] = []
    cur: List[complex] = []
    for r in rows[1:]:
        if not r:
            if cur:
                blocks.append(cur)
                cur = []
            continue
        if len(r) != 2:
            raise ValueError(f"{filename}: expected 're im' per coordinate line, got {' '.join(r)!r}")
        cur.append(complex(float(r[0]), float(r[1])))
    if cur:
        blocks.append(cur)
    if len(blocks) != count:
        raise ValueError(f"{filename}: header says {count} solutions, found {len(blocks)}")
    if count and len({len(b) for b in blocks}) != 1:
        raise ValueError(f"{filename}: solutions of different lengths")
    return np.array(blocks, dtype=np.complex128).reshape(count, -1 if count else 0)


def write_parameters(filename: PathLike, p: Sequence[complex]) -> None:
    """write_parameters(filename, p), src/utils.jl:109-114."""
    p = np.asarray(p, dtype=np.complex128).ravel()
    Path(filename).write_text("\n".join([str(len(p)), ""] + [_pair(z) for z in p]) + "\n")


def read_parameters(filename: PathLike) -> np.ndarray:
    """read_parameters(filename), src/utils.jl:128-134."""
    rows = [r for r in _rows(filename) if r]
    if not rows or len(rows[0]) != 1:
        raise ValueError(f"{filename}: first line must hold the number of parameters")
    n = int(float(rows[0][0]))
    if len(rows) - 1 < n:
        raise ValueError(f"{filename}: header says {n} values, found {len(rows) - 1}")
    return np.array([complex(float(r[0]), float(r[1])) for r in rows[1:n + 1]], dtype=np.complex128)


def read_seed_directory(directory: PathLike) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(start_parameters, start_solutions, target_parameters) of a seed directory."""
    d = Path(directory)
    return (read_parameters(d / "start_parameters.txt"), read_solutions(d / "start_solutions.txt"),
            read_parameters(d / "target_parameters.txt"))


def write_seed_directory(directory: PathLike, start_parameters, start_solutions, target_parameters) -> None:
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    write_parameters(d / "start_parameters.txt", start_parameters)
    write_solutions(d / "start_solutions.txt", start_solutions)
    write_parameters(d / "target_parameters.txt", target_parameters)
