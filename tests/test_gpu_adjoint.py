"""GPU parity of the topology-fixed window and its adjoint (extension).

* The window forward and the reverse pass on the GPU equal the oracle's bit
  for bit, for several checkpoint spacings (recomputed segments, a partial
  last segment) and for a batch of members with different parameters.
* A torch.autograd.Function over the C ABI (forward: sevc_window_forward,
  backward: sevc_adjoint_window) gives gradients of a torch-side smooth
  area-histogram loss that agree with differences of the GPU forward to
  1e-6 relative."""
import numpy as np
import pytest
import torch

import sevc
from _util import Oracle, diff_fields

pytestmark = pytest.mark.gpu


def _mesh(n=300, seed=4, gb_eps=0.0, **kw):
    pos, rid = sevc.voronoi(n, seed)
    rng = np.random.default_rng(seed)
    o = sevc.default_options()
    o.GB_eps = gb_eps
    o.GB_scaling = 1.3
    o.mobility = 0.8
    o.dt = 2e-5
    o.n_grains = n
    o.velocity_model = sevc.VEL_ANALYTIC
    o.events = 0
    o.do_nucleation = 0
    for k, v in kw.items():
        setattr(o, k, v)
    return o, sevc.build_mesh(o, pos, rid, rng.random(n), 0.5 + 2.0 * rng.random(n), n)


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("ckpt", [0, 1, 7, 25])
def test_window_and_adjoint_bitwise(ckpt):
    o, mesh = _mesh(gb_eps=0.3 if ckpt == 7 else 0.0)
    steps = 60
    print(f"  window {steps} steps, checkpoint every {ckpt}", flush=True)
    st, orc = sevc.State(o, mesh), Oracle(o, mesh)
    st.window_forward(steps, ckpt)
    assert orc.window_forward(steps) == 0
    assert diff_fields(st.get_mesh(), orc.get_mesh(), ["pos", "vel", "g_area", "b_energy"]) == {}
    rng = np.random.default_rng(ckpt)
    w = rng.standard_normal(2 * mesh.nv)
    (dp, ds, dgb, dmu), = st.adjoint(w)
    op, os_, ogb, omu = orc.adjoint(w)
    assert _same(dp, op) and _same(ds, os_)
    assert _same(dgb, ogb) and _same(dmu, omu), (dgb, ogb, dmu, omu)
    # a second reverse pass over the same window recomputes every segment
    (dp2, ds2, dgb2, dmu2), = st.adjoint(w)
    assert _same(dp2, op) and _same(ds2, os_) and dgb2 == ogb and dmu2 == omu
    st.close()


def test_adjoint_batch_members():
    o1, m1 = _mesh(n=200, seed=1)
    o2, m2 = _mesh(n=350, seed=2, mobility=1.2, GB_scaling=0.9, dt=1e-5)
    o3, m3 = _mesh(n=120, seed=3, gb_eps=0.2)
    st = sevc.State([o1, o2, o3], [m1, m2, m3])
    orcs = [Oracle(o, m) for o, m in ((o1, m1), (o2, m2), (o3, m3))]
    st.window_forward(45, 10)
    rng = np.random.default_rng(9)
    ws = [rng.standard_normal(2 * m.nv) for m in (m1, m2, m3)]
    got = st.adjoint(ws)
    for k, orc in enumerate(orcs):
        assert orc.window_forward(45) == 0
        assert diff_fields(st.get_mesh(k), orc.get_mesh(), ["pos", "vel"]) == {}, k
        ref = orc.adjoint(ws[k])
        for a, b in zip(got[k], ref):
            assert _same(a, b), k
    st.close()


def test_adjoint_needs_a_window():
    o, mesh = _mesh(n=100, seed=5)
    st = sevc.State(o, mesh)
    with pytest.raises(sevc.SevcError) as e:
        st.adjoint(np.zeros(2 * mesh.nv))
    assert e.value.code == sevc.E_STATE
    st.window_forward(5)
    st.step(1)                       # a step may change the topology
    with pytest.raises(sevc.SevcError):
        st.adjoint(np.zeros(2 * mesh.nv))
    st.close()


# ---------------------------------------------------------------------
# torch wrapper (test side): gradients through a torch loss
# ---------------------------------------------------------------------

class Window(torch.autograd.Function):
    """x_N = window(x_0, SE, GB_scaling, mobility) on one member."""

    @staticmethod
    def forward(ctx, pos0, se, gb, mu, state, steps, ckpt):
        state.set_params(0, pos=pos0.detach().cpu().numpy(), g_SE=se.detach().cpu().numpy(),
                         gb_scaling=float(gb), mobility=float(mu))
        state.window_forward(steps, ckpt)
        ctx.state = state
        return torch.from_numpy(state.get_mesh(0)["pos"].copy())

    @staticmethod
    def backward(ctx, g):
        (dp, ds, dgb, dmu), = ctx.state.adjoint(g.detach().cpu().numpy())
        return (torch.from_numpy(dp), torch.from_numpy(ds), torch.tensor(dgb, dtype=torch.float64),
                torch.tensor(dmu, dtype=torch.float64), None, None, None)


def _torch_hist_loss(pos, lists, mask, en, D, centres, width, target):
    """Smooth histogram of |A_g| (shoelace of the unwrapped list polygons)."""
    P = pos.reshape(-1, 2)
    X, Y = P[lists, 0], P[lists, 1]

    def wrap(d):
        return torch.where(d > D / 2, d - D, torch.where(d < -D / 2, D + d, d))
    dx = torch.where(mask[:, 1:], wrap(X[:, 1:] - X[:, :-1]), torch.zeros_like(X[:, 1:]))
    dy = torch.where(mask[:, 1:], wrap(Y[:, 1:] - Y[:, :-1]), torch.zeros_like(Y[:, 1:]))
    z = torch.zeros_like(X[:, :1])
    ux = torch.cat([z, torch.cumsum(dx, 1)], 1)
    uy = torch.cat([z, torch.cumsum(dy, 1)], 1)
    # closing term uses U_0 = 0, so the shifted arrays end with zeros
    uxn = torch.cat([ux[:, 1:], z], 1) * torch.cat([mask[:, 1:], mask[:, :1] & False], 1)
    uyn = torch.cat([uy[:, 1:], z], 1) * torch.cat([mask[:, 1:], mask[:, :1] & False], 1)
    A = 0.5 * torch.where(mask, ux * uyn - uy * uxn, torch.zeros_like(ux)).sum(1)
    a = torch.abs(A[en])
    zz = (a[:, None] - centres[None, :]) / width
    h = torch.exp(-0.5 * zz * zz).sum(0) / a.numel()
    return ((h - target) ** 2).sum()


def test_torch_function_gradients_match_differences():
    o, mesh = _mesh(n=250, seed=7)
    steps, ckpt = 40, 9
    st = sevc.State(o, mesh)
    gl = mesh.grain_lists()
    n = mesh["g_vlen"].astype(np.int64)
    en = (mesh["g_enabled"] == 1) & (n >= 3)
    mask = (np.arange(gl.shape[1])[None, :] < n[:, None]) & en[:, None]
    lists = torch.from_numpy(np.where(mask, gl, 0).astype(np.int64))
    mask_t, en_t = torch.from_numpy(mask), torch.from_numpy(en)
    from sevc.calib import grain_areas
    am = float(np.abs(grain_areas(mesh)).mean())
    centres = torch.from_numpy(am * (np.arange(20) + 0.5) / 10.0)
    target = torch.full((20,), 0.3, dtype=torch.float64)

    def loss_of(pos0, se, gb, mu):
        xN = Window.apply(pos0, se, gb, mu, st, steps, ckpt)
        return _torch_hist_loss(xN, lists, mask_t, en_t, mesh.D, centres, 0.1 * am, target)

    pos0 = torch.tensor(mesh["pos"], dtype=torch.float64, requires_grad=True)
    se = torch.tensor(mesh["g_SE"], dtype=torch.float64, requires_grad=True)
    gb = torch.tensor(o.GB_scaling, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(o.mobility, dtype=torch.float64, requires_grad=True)
    L = loss_of(pos0, se, gb, mu)
    L.backward()

    def fd(make, h):
        def f(e):
            with torch.no_grad():
                return float(loss_of(*make(e)))
        c1 = (f(h) - f(-h)) / (2 * h)
        c2 = (f(h / 2) - f(-h / 2)) / h
        return (4 * c2 - c1) / 3

    base = (pos0.detach(), se.detach(), gb.detach(), mu.detach())
    rng = np.random.default_rng(2)
    checks = []
    for c in rng.choice(np.nonzero(np.repeat(mesh["v_enabled"], 2))[0], 8, replace=False):
        def mk(e, c=c):
            p = base[0].clone()
            p[c] += e
            return (p,) + base[1:]
        checks.append((f"pos[{c}]", pos0.grad[c].item(), fd(mk, 1e-6)))
    for g in rng.choice(np.nonzero(en)[0], 4, replace=False):
        def mk(e, g=g):
            s = base[1].clone()
            s[g] += e
            return (base[0], s) + base[2:]
        checks.append((f"SE[{g}]", se.grad[g].item(), fd(mk, 1e-3)))
    checks.append(("gb", gb.grad.item(), fd(lambda e: (base[0], base[1], base[2] + e, base[3]), 1e-4)))
    checks.append(("mu", mu.grad.item(), fd(lambda e: (base[0], base[1], base[2], base[3] + e), 1e-4)))
    for name, ad, f in checks:
        assert abs(ad - f) <= 1e-6 * abs(f) + 1e-12, (name, ad, f)
    st.close()


def test_device_area_histogram_matches_numpy():
    """sevc_area_histogram (device) against sevc.calib (numpy): histogram,
    loss and dL/dx at the window's end; the device seed then drives the
    reverse pass exactly as the same gradient passed from the host."""
    from sevc.calib import AreaHistogram, grain_areas
    o1, m1 = _mesh(n=400, seed=11)
    o2, m2 = _mesh(n=250, seed=12, mobility=1.1)
    st = sevc.State([o1, o2], [m1, m2])
    st.window_forward(20, 6)
    ranges, hs = [], []
    for m in (m1, m2):
        am = float(np.abs(grain_areas(m)).mean())
        ranges.append((0.0, 3.0 * am, 0.1 * am))
        hs.append(AreaHistogram(20, 0.0, 3.0 * am, 0.1 * am))
    targets = np.random.default_rng(0).random((2, 20)) * 0.5
    hist, loss, grads = st.area_histogram(20, ranges, targets, seed=True, want_grad=True)
    for k, (h, m) in enumerate(zip(hs, (m1, m2))):
        P = st.positions(k)
        Lr, gr = h.loss_and_grad(m, P, targets[k])
        assert np.allclose(hist[k], h(m, P), rtol=1e-12, atol=1e-15), k
        assert abs(loss[k] - Lr) <= 1e-12 * abs(Lr), (k, loss[k], Lr)
        assert np.allclose(grads[k], gr, rtol=1e-10, atol=1e-12 * np.abs(gr).max()), k
    seeded = st.adjoint(None)
    host = st.adjoint(grads)
    for a, b in zip(seeded, host):
        for x, y in zip(a, b):
            assert _same(x, y)
    st.close()
