"""The 200-bit endpoint refiner used by the per-configuration endpoint checks: it reaches
exact zeros known in closed form, and it agrees with full 200-bit Newton (Jacobian in
200 bits too) on a dense cubic system."""
import mpmath
import numpy as np

from hcgpu.polys import System, variables
from refine import refine


def test_refine_reaches_closed_form_zero():
    (x, y, z), _ = variables(3)
    # zeros: x = sqrt(2), y = x + 1/3, z = x y - 5
    F = System([x * x - 2, 3 * y - 3 * x - 1, z - x * y + 5], 3)
    x0 = np.array([1.4142, 1.7475, -2.53])
    R, last = refine(F, x0[None, :])
    mpmath.mp.prec = 200
    s2 = mpmath.sqrt(2)
    exact = [s2, s2 + mpmath.mpf(1) / 3, s2 * (s2 + mpmath.mpf(1) / 3) - 5]
    for i in range(3):
        assert abs(complex(exact[i]) - R[0, i]) <= 1e-15 * abs(complex(exact[i]))
    assert last[0] < 1e-40


def test_refine_agrees_with_full_newton():
    from hcgpu.configs import dense_system
    from test_oracle_known_answers import _newton_mp
    import oracle
    from hcgpu.homotopies import total_degree
    from hcgpu.solve import make_options
    F = dense_system(4, 3, seed=100)
    H, S = total_degree(F, seed=101)
    o = oracle.track_batch(H, S[:12], make_options())
    R, last = refine(F, o.solution)
    for k in range(12):
        xr = _newton_mp(F, o.solution[k])
        assert np.abs(R[k] - xr).max() <= 1e-15 * max(1.0, np.abs(xr).max())
    assert (last < 1e-40).all()
