"""Calibration losses on the grain-area distribution.

The adjoint window (State.window_forward / State.adjoint) takes dL/dx at
the window's end; this module supplies a smooth loss of the grain areas and
its gradient with respect to the vertex positions, vectorised over the
padded grain lists of a host mesh.

Areas are the signed shoelace areas of the periodic polygons (each vertex
unwrapped against its list predecessor), the same quantity the time step
computes; the loss sees |A_g|.  Within a window the topology is fixed, so
the list structure is prepared once per mesh and reused.
"""
import numpy as np


def _wrap(d, D):
    """Minimum-image difference of geometry.h:147-152 (|d| <= D/2 kept)."""
    return np.where(d > D / 2, d - D, np.where(d < -D / 2, D + d, d))


class _Lists:
    """Padded grain lists of a mesh trimmed to the longest list, with masks
    and the list predecessor / successor of every entry."""

    def __init__(self, mesh):
        gl = mesh.grain_lists()
        n = mesh["g_vlen"].astype(np.int64)
        self.en = (mesh["g_enabled"] == 1) & (n >= 3)
        S = max(3, int(n[self.en].max()) if self.en.any() else 3)
        k = np.arange(S)[None, :]
        self.mask = (k < n[:, None]) & self.en[:, None]
        self.idx = np.where(self.mask, gl[:, :S], 0)
        nn = np.maximum(n, 1)[:, None]
        self.prv = np.take_along_axis(self.idx, (k - 1) % nn, axis=1)
        self.nxt = np.take_along_axis(self.idx, (k + 1) % nn, axis=1)
        self.last = np.where(self.en, n - 1, 0)
        self.flat = self.idx[self.mask]
        self.D = mesh.D
        self.nv = mesh.nv


_CACHE = {}


def _lists(mesh):
    key = id(mesh)
    hit = _CACHE.get(key)
    if hit is None or hit[0] is not mesh:
        hit = (mesh, _Lists(mesh))
        _CACHE[key] = hit
    return hit[1]


def grain_areas(mesh, pos=None):
    """Signed area of every grain slot (0 for disabled grains)."""
    s = _lists(mesh)
    D = s.D
    P = mesh.xy() if pos is None else np.asarray(pos, dtype=np.float64).reshape(-1, 2)
    X, Y = P[s.idx, 0], P[s.idx, 1]
    m1 = s.mask[:, 1:]
    dx = np.where(m1, _wrap(X[:, 1:] - X[:, :-1], D), 0.0)
    dy = np.where(m1, _wrap(Y[:, 1:] - Y[:, :-1], D), 0.0)
    ng = X.shape[0]
    ux = np.concatenate([np.zeros((ng, 1)), np.cumsum(dx, axis=1)], axis=1)
    uy = np.concatenate([np.zeros((ng, 1)), np.cumsum(dy, axis=1)], axis=1)
    # U_{k+1} for k < n-1, and U_0 = 0 for the closing term
    uxn = np.concatenate([ux[:, 1:], np.zeros((ng, 1))], axis=1) * np.concatenate([m1, np.zeros((ng, 1), bool)], 1)
    uyn = np.concatenate([uy[:, 1:], np.zeros((ng, 1))], axis=1) * np.concatenate([m1, np.zeros((ng, 1), bool)], 1)
    term = np.where(s.mask, ux * uyn - uy * uxn, 0.0)
    return np.where(s.en, 0.5 * term.sum(axis=1), 0.0)


def area_gradient(mesh, pos, w):
    """sum_g w_g dA_g/dx for per-grain weights w: dA/dx_k = (dy_prev + dy_next)/2,
    dA/dy_k = -(dx_prev + dx_next)/2 with wrapped list differences.
    Returns shape (n_vertex_slots, 2)."""
    s = _lists(mesh)
    D = s.D
    P = np.asarray(pos, dtype=np.float64).reshape(-1, 2)
    X, Y = P[s.idx, 0], P[s.idx, 1]
    dpx, dpy = _wrap(X - P[s.prv, 0], D), _wrap(Y - P[s.prv, 1], D)
    dnx, dny = _wrap(P[s.nxt, 0] - X, D), _wrap(P[s.nxt, 1] - Y, D)
    ww = np.asarray(w, dtype=np.float64)[:, None]
    gx = (0.5 * ww * (dpy + dny))[s.mask]
    gy = (-0.5 * ww * (dpx + dnx))[s.mask]
    g = np.empty((s.nv, 2))
    g[:, 0] = np.bincount(s.flat, weights=gx, minlength=s.nv)
    g[:, 1] = np.bincount(s.flat, weights=gy, minlength=s.nv)
    return g


class AreaHistogram:
    """Smooth histogram of |A_g| over the enabled grains:
    h_k = (1/G) sum_g exp(-((|A_g| - c_k)/w)^2 / 2), c_k the centres of
    n_bins equal bins on [a_min, a_max], w = width."""

    def __init__(self, n_bins, a_min, a_max, width):
        self.c = a_min + (np.arange(n_bins) + 0.5) * (a_max - a_min) / n_bins
        self.w = float(width)

    def __call__(self, mesh, pos=None):
        A = grain_areas(mesh, pos)
        a = np.abs(A[_lists(mesh).en])
        z = (a[:, None] - self.c[None, :]) / self.w
        return np.exp(-0.5 * z * z).sum(axis=0) / max(1, a.size)

    def loss_and_grad(self, mesh, pos, target):
        """L = sum_k (h_k - t_k)^2 and dL/dx (2 per vertex slot)."""
        P = mesh.xy() if pos is None else np.asarray(pos, dtype=np.float64).reshape(-1, 2)
        en = _lists(mesh).en
        A = grain_areas(mesh, P)
        G = max(1, int(en.sum()))
        a = np.abs(A[en])
        z = (a[:, None] - self.c[None, :]) / self.w
        e = np.exp(-0.5 * z * z)
        h = e.sum(axis=0) / G
        r = h - np.asarray(target, dtype=np.float64)
        L = float(np.sum(r * r))
        wg = np.zeros(len(A))
        wg[en] = (2.0 * r[None, :] * e * (-z / self.w)).sum(axis=1) / G * np.where(A[en] < 0, -1.0, 1.0)
        return L, area_gradient(mesh, P, wg).reshape(-1)
