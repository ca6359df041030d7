"""Systems of the reference's own tests (test/endgame_test.jl, test/endgame_tracker_test.jl,
test/tracker_test.jl, test/test_systems.jl), written with the polynomial front end."""
import numpy as np

from hcgpu.polys import Poly, System, variables


def beyond_polyhedral():
    (x, y), _ = variables(2)
    return System([2.3 * x**2 + 1.2 * y**2 + 3 * x - 2 * y + 3, 2.3 * x**2 + 1.2 * y**2 + 5 * x + 2 * y - 5], 2)


def cyclic(n):
    z, _ = variables(n)
    eqs = []
    for m in range(n - 1):
        s = Poly.const(0, n)
        for j in range(n):
            p = Poly.const(1, n)
            for k in range(j, j + m + 1):
                p = p * z[k % n]
            s = s + p
        eqs.append(s)
    p = Poly.const(1, n)
    for v in z:
        p = p * v
    eqs.append(p - 1)
    return System(eqs, n)


def hyperbolic():
    (x, z), _ = variables(2)
    y = 1
    return System([0.75 * x**4 + 1.5 * x**2 * y**2 - 2.5 * x**2 * z**2 + 0.75 * y**4 - 2.5 * y**2 * z**2 + 0.75 * z**4,
                   10 * x**2 * z + 10 * y**2 * z - 6 * z**3], 2)


def singular1():
    (x, y), _ = variables(2)
    z = 1
    return System([x**2 + 2 * y**2 + 2j * y * z,
                   (18 + 3j) * x * y + 7j * y**2 - (3 - 18j) * x * z - 14 * y * z - 7j * z**2], 2)


def wilkinson(d):
    (x,), _ = variables(1)
    f = Poly.const(1, 1)
    for i in range(1, d + 1):
        f = f * (x - i)
    return System([f], 1)


def winding_family(d):
    a = [0.257, -0.139, -1.73, -0.199, 1.79, -1.32]
    (x, y), _ = variables(2)
    f1 = (a[0] * x**d + a[1] * y) * (a[2] * x + a[3] * y) + 1
    f2 = (a[0] * x**d + a[1] * y) * (a[4] * x + a[5] * y) + 1
    return System([f1, f2], 2)


def mohab():
    (x, z, y), _ = variables(3)  # variable order [x, z, y] as in the reference test
    return System([
        -9091098778555951517 * x**3 * y**4 * z**2 + 5958442613080401626 * y**2 * z**7
        + 17596733865548170996 * x**2 * z**6 - 17979170986378486474 * x * y * z**6
        - 2382961149475678300 * x**4 * y**3 - 15412758154771986214 * x * y**3 * z**3 + 133,
        -10798198881812549632 * x**6 * y**3 * z - 11318272225454111450 * x * y**9
        - 14291416869306766841 * y**9 * z - 5851790090514210599 * y**2 * z**8
        + 15067068695242799727 * x**2 * y**3 * z**4 + 7716112995720175148 * x**3 * y * z**3 + 171,
        13005416239846485183 * x**7 * y**3 + 4144861898662531651 * x**5 * z**4 - 8026818640767362673 * x**6
        - 6882178109031199747 * x**2 * y**4 + 7240929562177127812 * x**2 * y**3 * z
        + 5384944853425480296 * x * y * z**4 + 88,
    ], 3)


MOHAB_GAMMA = -0.9132549847010242 + 0.4073884300256109j


def tracker_param_system():
    """F = [x^2 - a, x*y - a + b] with parameters (a, b) (test/tracker_test.jl:5-27)."""
    (x, y), (a, b) = variables(2, 2)
    return System([x**2 - a, x * y - a + b], 2, 2)
