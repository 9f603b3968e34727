"""Closed-form pins of the oracle's periodic geometry (geometry.h)."""
from fractions import Fraction

import numpy as np

from _util import oracle_lib


def test_wrap_and_mod_identities():
    L = oracle_lib()
    rng = np.random.default_rng(0)
    for D in (1.0, 40.0):
        a, b = rng.uniform(0, D, 5000), rng.uniform(0, D, 5000)
        w = np.array([L.orc_wrap_dist(x, y, D) for x, y in zip(a, b)])
        assert np.all(np.abs(w) <= D / 2)
        # minimum image: w - (a - b) is a multiple of D
        k = (w - (a - b)) / D
        assert np.allclose(k, np.round(k), atol=1e-12)
        # antisymmetry is exact
        w2 = np.array([L.orc_wrap_dist(y, x, D) for x, y in zip(a, b)])
        assert np.array_equal(w, -w2)
        m = np.array([L.orc_dom_mod(x, D) for x in np.concatenate([a - D, a + D])])
        assert np.all((m >= 0) & (m < D))


def test_mag_close_to_hypot():
    L = oracle_lib()
    rng = np.random.default_rng(1)
    x, y = rng.normal(size=20000) * 10.0 ** rng.integers(-8, 3, 20000), rng.normal(size=20000)
    m = np.array([L.orc_mag(a, b) for a, b in zip(x, y)])
    h = np.hypot(x, y)
    assert np.all(np.abs(m - h) <= 2 * np.spacing(h))
    assert L.orc_mag(0.0, 0.0) == 0.0
    assert L.orc_mag(3.0, 4.0) == 5.0


def _area(xs, ys, D=1.0):
    import ctypes as C
    L = oracle_lib()
    n = len(xs)
    X = (C.c_double * n)(*xs)
    Y = (C.c_double * n)(*ys)
    return L.orc_polygon_area(X, Y, n, D)


def test_polygon_area_exact_on_dyadic_points():
    # vertices on a 1/64 grid and less than D/2 apart (no wrapping): every
    # product and sum is exact in FP64
    rng = np.random.default_rng(2)
    for _ in range(200):
        n = int(rng.integers(3, 12))
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        xs = [Fraction(int(round(64 * (0.5 + 0.2 * np.cos(t)))), 64) for t in ang]
        ys = [Fraction(int(round(64 * (0.5 + 0.2 * np.sin(t)))), 64) for t in ang]
        exact = sum(xs[k] * ys[(k + 1) % n] - ys[k] * xs[(k + 1) % n] for k in range(n)) / 2
        got = _area([float(v) for v in xs], [float(v) for v in ys])
        # the oracle unwraps about vertex 0: translation does not change the area
        assert got == float(exact)


def test_polygon_area_across_the_periodic_boundary():
    # a unit square of side 0.25 straddling the corner of the torus
    xs = [0.875, 0.125, 0.125, 0.875]
    ys = [0.875, 0.875, 0.125, 0.125]
    assert _area(xs, ys) == 0.0625
    assert _area(xs[::-1], ys[::-1]) == -0.0625


def test_chunked_sum_order():
    import ctypes as C
    L = oracle_lib()
    rng = np.random.default_rng(3)
    v = rng.random(1000)
    for n in (1, 200, 256):
        arr = (C.c_double * n)(*v[:n])
        s = 0.0
        for x in v[:n]:
            s += x
        assert L.orc_sum_chunked(arr, None, n) == s
    n = 1000
    arr = (C.c_double * n)(*v)
    s = 0.0
    for c in range(0, n, 256):
        t = 0.0
        for x in v[c:c + 256]:
            t += x
        s += t
    assert L.orc_sum_chunked(arr, None, n) == s
