"""Writes cfg2_starts.npy: the 1,024 start solutions of the batched parameter-homotopy
workload (n = 10 sparse quadrics with coefficients affine in 16 parameters, at the start
parameters p0), produced by the oracle's total-degree solve of F(x; p0) with every
nonsingular endpoint kept and refined by Newton's method in 200-bit arithmetic.

Run from the repository root:  python tests/golden/make_cfg2_starts.py
"""
import sys
from pathlib import Path

import mpmath
import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "homotopycontinuation.jl_amd"))

import oracle  # noqa: E402
from hcgpu.configs import CFG2_SEEDS, cfg2_problem  # noqa: E402
from hcgpu.homotopies import total_degree  # noqa: E402
from hcgpu.solve import make_options  # noqa: E402


def refine(F, x, iters=6):
    mpmath.mp.prec = 200
    n = F.nvars
    z = [mpmath.mpc(complex(v)) for v in x]
    polys = [[(ex, mpmath.mpc(complex(c))) for ex, c in f.terms.items()] for f in F.polys]
    for _ in range(iters):
        fv, J = [], mpmath.matrix(n, n)
        for i, terms in enumerate(polys):
            s = mpmath.mpc(0)
            for ex, c in terms:
                m = c
                for v, e in zip(z, ex):
                    if e:
                        m *= v ** e
                s += m
                for j in range(n):
                    if ex[j]:
                        d = c * ex[j]
                        for k, (v, e) in enumerate(zip(z, ex)):
                            ee = e - 1 if k == j else e
                            if ee:
                                d *= v ** ee
                        J[i, j] += d
            fv.append(s)
        dx = mpmath.lu_solve(J, mpmath.matrix(fv))
        z = [z[i] - dx[i] for i in range(n)]
    return np.array([complex(v) for v in z])


def main():
    F, p0, _ = cfg2_problem()
    Fp = F.fix_parameters(p0)
    H, S = total_degree(Fp, seed=CFG2_SEEDS["gamma"])
    o = oracle.track_batch(H, S, make_options())
    keep = (o.code == 1) & (~o.singular)
    X = o.solution[keep]
    print(f"{len(S)} paths, {keep.sum()} nonsingular endpoints")
    R = np.stack([refine(Fp, x) for x in X])
    rel = np.abs(R - X).max(axis=1) / np.maximum(1, np.abs(R).max(axis=1))
    print(f"max relative change under 200-bit refinement: {rel.max():.2e}")
    np.save(HERE / "cfg2_starts.npy", R)


if __name__ == "__main__":
    main()
