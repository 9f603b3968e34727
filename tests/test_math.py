"""The shared arithmetic of include/sevc_math.h, compiled on the host.

Pins: the Random123 known-answer vectors of Philox4x32-10, and the Cephes
elementary functions against numpy/glibc within 3 ulp (a wrong coefficient
or reduction constant shows up as errors of millions of ulp)."""
import ctypes as C

import numpy as np
import pytest

from _util import oracle_lib


def _philox(ctr, key):
    L = oracle_lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.orc_philox(c, k, o)
    return [o[i] for i in range(4)]


@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, expect):
    assert _philox(ctr, key) == expect


def _ulps(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))


def test_atan2_accuracy_and_quadrants():
    L = oracle_lib()
    rng = np.random.default_rng(3)
    x, y = rng.uniform(-50, 50, 50000), rng.uniform(-50, 50, 50000)
    r = np.array([L.orc_atan2(b, a) for a, b in zip(x, y)])
    assert _ulps(r, np.arctan2(y, x)).max() <= 3
    for yy, xx in [(0.0, 1.0), (0.0, -1.0), (1.0, 0.0), (-1.0, 0.0), (1.0, 1.0), (-1.0, -1.0), (1e-300, -1.0)]:
        assert abs(L.orc_atan2(yy, xx) - np.arctan2(yy, xx)) <= 4e-16 * max(1.0, abs(np.arctan2(yy, xx)))


def test_sin_cos_accuracy():
    L = oracle_lib()
    rng = np.random.default_rng(4)
    t = np.concatenate([rng.uniform(-np.pi, np.pi, 40000), rng.uniform(-300, 300, 20000)])
    assert _ulps([L.orc_sin(v) for v in t], np.sin(t)).max() <= 3
    assert _ulps([L.orc_cos(v) for v in t], np.cos(t)).max() <= 3
    # the t_ext tangent cos(atan2(y,x)), sin(atan2(y,x)) is the unit vector
    x, y = rng.uniform(-1, 1, 5000), rng.uniform(-1, 1, 5000)
    ux = np.array([L.orc_cos(L.orc_atan2(b, a)) for a, b in zip(x, y)])
    uy = np.array([L.orc_sin(L.orc_atan2(b, a)) for a, b in zip(x, y)])
    r = np.hypot(x, y)
    assert np.abs(ux - x / r).max() < 2e-15 and np.abs(uy - y / r).max() < 2e-15


def test_exp_accuracy():
    L = oracle_lib()
    rng = np.random.default_rng(5)
    e = rng.uniform(-700, 700, 50000)
    assert _ulps([L.orc_exp(v) for v in e], np.exp(e)).max() <= 3
    assert L.orc_exp(0.0) == 1.0
    assert L.orc_exp(-1000.0) == 0.0
