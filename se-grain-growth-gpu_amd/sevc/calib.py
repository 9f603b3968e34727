"""Calibration losses on the grain-area distribution.

The adjoint window (State.window_forward / State.adjoint) takes dL/dx at
the window's end; this module supplies a smooth loss of the grain areas and
its gradient with respect to the vertex positions, vectorised over the
padded grain lists of a host mesh.

Areas are the signed shoelace areas of the periodic polygons (each vertex
unwrapped against its list predecessor), the same quantity the time step
computes; the loss sees |A_g|.
"""
import numpy as np


def _wrap(d, D):
    """Minimum-image difference of geometry.h:147-152 (|d| <= D/2 kept)."""
    return np.where(d > D / 2, d - D, np.where(d < -D / 2, D + d, d))


def _lists(mesh):
    gl = mesh.grain_lists()
    n = mesh["g_vlen"].astype(np.int64)
    en = (mesh["g_enabled"] == 1) & (n >= 3)
    S = gl.shape[1]
    k = np.arange(S)[None, :]
    mask = (k < n[:, None]) & en[:, None]
    idx = np.where(mask, gl, 0)
    return idx, mask, n, en


def grain_areas(mesh, pos=None):
    """Signed area of every grain slot (0 for disabled grains)."""
    D = mesh.D
    P = mesh.xy() if pos is None else np.asarray(pos, dtype=np.float64).reshape(-1, 2)
    idx, mask, n, en = _lists(mesh)
    X, Y = P[idx, 0], P[idx, 1]
    dx = np.where(mask[:, 1:], _wrap(X[:, 1:] - X[:, :-1], D), 0.0)
    dy = np.where(mask[:, 1:], _wrap(Y[:, 1:] - Y[:, :-1], D), 0.0)
    ux = np.concatenate([np.zeros((len(n), 1)), np.cumsum(dx, axis=1)], axis=1)
    uy = np.concatenate([np.zeros((len(n), 1)), np.cumsum(dy, axis=1)], axis=1)
    nxt = np.roll(np.arange(idx.shape[1]), -1)
    last = np.where(en, n - 1, 0)
    uxn, uyn = ux[:, nxt].copy(), uy[:, nxt].copy()
    rows = np.arange(len(n))
    uxn[rows, last], uyn[rows, last] = 0.0, 0.0        # closing term to U_0 = 0
    term = np.where(mask, ux * uyn - uy * uxn, 0.0)
    return np.where(en, 0.5 * term.sum(axis=1), 0.0)


def area_gradient(mesh, pos, w):
    """sum_g w_g dA_g/dx for per-grain weights w: dA/dx_k = (dy_prev + dy_next)/2,
    dA/dy_k = -(dx_prev + dx_next)/2 with wrapped list differences."""
    D = mesh.D
    P = np.asarray(pos, dtype=np.float64).reshape(-1, 2)
    idx, mask, n, en = _lists(mesh)
    S = idx.shape[1]
    k = np.arange(S)[None, :]
    nn = np.maximum(n, 1)[:, None]
    prv = np.take_along_axis(idx, (k - 1) % nn, axis=1)
    nxt = np.take_along_axis(idx, (k + 1) % nn, axis=1)
    X, Y = P[idx, 0], P[idx, 1]
    dpx, dpy = _wrap(X - P[prv, 0], D), _wrap(Y - P[prv, 1], D)
    dnx, dny = _wrap(P[nxt, 0] - X, D), _wrap(P[nxt, 1] - Y, D)
    ww = np.where(mask, np.asarray(w, dtype=np.float64)[:, None], 0.0)
    g = np.zeros_like(P)
    np.add.at(g[:, 0], idx[mask], (0.5 * ww * (dpy + dny))[mask])
    np.add.at(g[:, 1], idx[mask], (-0.5 * ww * (dpx + dnx))[mask])
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
        en = (mesh["g_enabled"] == 1) & (mesh["g_vlen"] >= 3)
        a = np.abs(A[en])
        z = (a[:, None] - self.c[None, :]) / self.w
        return np.exp(-0.5 * z * z).sum(axis=0) / max(1, a.size)

    def loss_and_grad(self, mesh, pos, target):
        """L = sum_k (h_k - t_k)^2 and dL/dx (2 per vertex slot)."""
        P = mesh.xy() if pos is None else np.asarray(pos, dtype=np.float64).reshape(-1, 2)
        A = grain_areas(mesh, P)
        en = (mesh["g_enabled"] == 1) & (mesh["g_vlen"] >= 3)
        G = max(1, int(en.sum()))
        a = np.abs(A)
        z = (a[:, None] - self.c[None, :]) / self.w
        e = np.where(en[:, None], np.exp(-0.5 * z * z), 0.0)
        h = e.sum(axis=0) / G
        r = h - np.asarray(target, dtype=np.float64)
        L = float(np.sum(r * r))
        dL_da = (2.0 * r[None, :] * e * (-z / self.w)).sum(axis=1) / G      # d/d|A_g|
        wg = np.where(en, dL_da * np.where(A < 0, -1.0, 1.0), 0.0)
        return L, area_gradient(mesh, P, wg).reshape(-1)
