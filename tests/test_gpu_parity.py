"""GPU parity: the HIP time step against the serial FP64 oracle.

Integer topology and every FP64 field must match bit for bit (the oracle
and the kernels evaluate the same operations in the same order with no
fused multiply-add)."""
import numpy as np
import pytest

import sevc
from _util import Oracle, diff_fields, ref2k_options

pytestmark = pytest.mark.gpu


def _pair(opts, mesh):
    return sevc.State(opts, mesh), Oracle(opts, mesh)


def test_selftest_math_bits():
    rng = np.random.default_rng(0)
    n = 20000
    x = rng.uniform(-4, 4, n)
    y = rng.uniform(-4, 4, n)
    from _util import oracle_lib
    L = oracle_lib()
    cases = {0: lambda a, b: L.orc_atan2(a, b), 1: lambda a, b: L.orc_sin(a), 2: lambda a, b: L.orc_cos(a),
             3: lambda a, b: L.orc_exp(a), 4: lambda a, b: a / b, 6: lambda a, b: L.orc_mag(a, b),
             7: lambda a, b: L.orc_wrap_dist(a, b, 1.0)}
    for op, f in cases.items():
        dev = sevc.selftest_math(op, x, y)
        host = np.array([f(a, b) for a, b in zip(x, y)])
        assert np.array_equal(dev.view(np.uint64), host.view(np.uint64)), f"op {op} differs"
    xs = np.abs(x) * 10.0 ** rng.integers(-20, 20, n)
    dev = sevc.selftest_math(5, xs, y)
    assert np.array_equal(dev.view(np.uint64), np.sqrt(xs).view(np.uint64))


def test_step0_velocities_bitwise():
    o = ref2k_options(events=0)
    mesh = sevc.load_files(o)
    st, orc = _pair(o, mesh)
    st.step(1)
    orc.step(1)
    a, b = st.get_mesh(), orc.get_mesh()
    assert diff_fields(a, b) == {}


@pytest.mark.parametrize("steps", [50, 400])
def test_steps_with_events_bitwise(steps):
    o = ref2k_options()
    mesh = sevc.load_files(o)
    st, orc = _pair(o, mesh)
    s1 = st.step(steps)
    rc, s2 = orc.step(steps)
    assert rc == 0 and s1.status == 0
    a, b = st.get_mesh(), orc.get_mesh()
    assert diff_fields(a, b) == {}
    for k in ("n", "n_grains", "n_vertices", "n_flips", "n_removals", "n_two_grain", "n_vertex_steps"):
        assert getattr(s1, k) == getattr(s2, k), k
