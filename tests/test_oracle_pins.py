"""Pins of the oracle's vote / conflict rounds and of its nucleation that do
not come from the oracle itself.

* Vote rounds (vertices_vote / count_boundaries_votes / unlabel until no
  change, topological.h:362-457, polling_system :549-578): with distinct
  extinction times the fixpoint is the greedy matching of the candidate
  edges taken in increasing t_ext (an edge is kept when neither end already
  holds a kept edge), computed here by that sequential algorithm.  The
  result is an independent set of edges and a fixpoint (another round
  changes nothing).  A tie at a vertex goes to the later boundary slot
  (`<=`, :371), checked on a hand-built case.
* Conflict filter (topological.h:83-155): a candidate survives iff it beats
  (smaller t_ext, then smaller id) every candidate joined to it by a
  non-candidate edge; afterwards no non-candidate edge has candidates at
  both ends.
* Nucleation energetics (generate_grain_for_nucleation,
  nucleation2.cu:116-212): Delta E = E1 - E0 in closed form from the three
  angles theta_i between the boundaries at the site: the new vertices sit at
  distance r on AD, AE, AF (diagram :83-92), the three small triangles have
  areas r^2 sin(theta_i) / 2, the new boundaries lengths 2 r sin(theta_i/2),
  with g1 between AF and AD, g2 between AD and AE, g3 between AE and AF.
  Delta^2 E = Delta E(r) - Delta E(0.99 r) takes the closed form's sign.
* Insertion (nucleate, nucleation2.cu:474-676, which places the nucleus
  at 0.99 r, :526-527): positions, rewiring, grain lists and areas of the
  diagram, counter-clockwise boundary order at the three new vertices;
  check_data, Euler and total area.
* Statistics: Metropolis acceptance (:414-440) happens with probability
  exp(-MC_k dE) on the sites with dE >= 0 and Delta^2 E < 0 (always for
  dE < 0, never otherwise) and the random site is uniform over alive
  vertices; in the rate mode grain g draws with probability rate * SE_g per
  step at a uniformly chosen vertex of its list.

Each check is a function returning None or a failure message, so that the
same checks can be run against deliberately broken copies of the oracle
(compiled with -DORC_MUTANT=k, one planted mistake each): every planted
mistake must fail at least one check."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest
from scipy import stats

import sevc
from sevc._abi import HostMesh, Mesh, Options, StepStats
from _util import ORACLE_SO, REPO, oracle_lib

P = C.POINTER


def _bind(L):
    L.orc_create.restype = C.c_void_p
    L.orc_create.argtypes = [P(Options), P(Mesh)]
    L.orc_destroy.argtypes = [C.c_void_p]
    L.orc_step.argtypes = [C.c_void_p, C.c_int32, P(StepStats)]
    L.orc_get_mesh.argtypes = [C.c_void_p, P(Mesh)]
    L.orc_check_mesh.argtypes = [P(Mesh), P(C.c_double)]
    L.orc_poll.argtypes = [C.c_void_p, P(C.c_double), P(C.c_uint8), P(C.c_int32), C.c_int32]
    L.orc_nucl_geometry.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_int32, P(C.c_double),
                                    P(C.c_double), P(C.c_int32)]
    L.orc_nucleate_vertex.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double, P(C.c_int32)]
    L.orc_mc_trial.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_double), P(C.c_double)]
    L.orc_rate_draws.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int32), C.c_int32]
    return L


class Orc:
    def __init__(self, L, o, h):
        self.L, self.h0 = L, h
        self.p = L.orc_create(C.byref(o), C.byref(h.struct))
        assert self.p

    def mesh(self):
        t = self.h0
        out = HostMesh(t.nv, t.ne, t.ng, t.S, t.D)
        assert self.L.orc_get_mesh(self.p, C.byref(out.struct)) == 0
        return out

    def poll(self, t_ext, cand, stages):
        t = np.ascontiguousarray(t_ext, dtype=np.float64)
        c = np.ascontiguousarray(cand, dtype=np.uint8).copy()
        conf = np.zeros(t.size, dtype=np.int32)
        self.L.orc_poll(self.p, t.ctypes.data_as(P(C.c_double)), c.ctypes.data_as(P(C.c_uint8)),
                        conf.ctypes.data_as(P(C.c_int32)), int(stages))
        return c.astype(bool), conf

    def geometry(self, A, r, ori, do_scale=0):
        xy = np.zeros(6)
        de = C.c_double()
        ids = np.zeros(9, dtype=np.int32)
        rc = self.L.orc_nucl_geometry(self.p, int(A), float(r), float(ori), int(do_scale),
                                      xy.ctypes.data_as(P(C.c_double)), C.byref(de),
                                      ids.ctypes.data_as(P(C.c_int32)))
        return rc, xy.reshape(3, 2), de.value, ids

    def nucleate(self, A, r, ori):
        slots = np.zeros(6, dtype=np.int32)
        rc = self.L.orc_nucleate_vertex(self.p, int(A), float(r), float(ori), slots.ctypes.data_as(P(C.c_int32)))
        return rc, slots

    def mc_trial(self, step):
        site, de, nf = C.c_int32(), C.c_double(), C.c_double()
        acc = self.L.orc_mc_trial(self.p, int(step), C.byref(site), C.byref(de), C.byref(nf))
        return acc, site.value, de.value, nf.value

    def rate_draws(self, step, cap):
        g = np.zeros(cap, dtype=np.int32)
        v = np.zeros(cap, dtype=np.int32)
        n = self.L.orc_rate_draws(self.p, int(step), g.ctypes.data_as(P(C.c_int32)), v.ctypes.data_as(P(C.c_int32)),
                                  cap)
        return g[:min(n, cap)], v[:min(n, cap)]

    def close(self):
        if self.p:
            self.L.orc_destroy(self.p)
            self.p = None

    def __del__(self):
        self.close()


def _mesh(n, seed, **kw):
    pos, rid = sevc.voronoi(n, seed)
    rng = np.random.default_rng(seed)
    o = sevc.default_options()
    o.GB_eps = 0.0
    o.n_grains = n
    for k, v in kw.items():
        setattr(o, k, v)
    return o, sevc.build_mesh(o, pos, rid, rng.random(n), rng.random(n), n)


def _wrap(d, D):
    return d - D * np.round(d / D)


# ---------------------------------------------------------------------
# vote rounds and conflict filter
# ---------------------------------------------------------------------

def greedy_matching(cand, t, ends):
    """Candidates taken in increasing t_ext; an edge is kept when neither
    of its ends already holds a kept edge."""
    used = set()
    keep = np.zeros(cand.size, dtype=bool)
    for e in sorted(np.nonzero(cand)[0], key=lambda e: t[e]):
        a, b = ends[e]
        if a in used or b in used:
            continue
        keep[e] = True
        used.update((a, b))
    return keep


def conflict_rule(S, t, ends, vb):
    """A kept edge survives iff it beats, by (t_ext, id), every kept edge
    joined to it by a non-candidate edge."""
    out = S.copy()
    at = {}
    for e in np.nonzero(S)[0]:
        for v in ends[e]:
            at[v] = e
    for e in np.nonzero(S)[0]:
        for p in ends[e]:
            for X in vb[p]:
                if X == e or S[X]:
                    continue
                q = ends[X][1] if ends[X][0] == p else ends[X][0]
                e2 = at.get(q)
                if e2 is None or e2 == e:
                    continue
                if (t[e2], e2) < (t[e], e):
                    out[e] = False
    return out


def _topology(h):
    ends = h["b_vtx"].reshape(-1, 2)
    vb = h["v_bnd"].reshape(-1, 3)
    return [tuple(x) for x in ends], [list(x) for x in vb]


def check_polling(L):
    for n, seed in ((12, 3), (20, 5), (40, 7)):
        o, h = _mesh(n, seed)
        orc = Orc(L, o, h)
        ends, vb = _topology(h)
        rng = np.random.default_rng(100 + n)
        for trial in range(60):
            for p in (0.3, 0.6, 0.95):
                t = rng.uniform(0.0, 1.0, h.ne)
                cand = rng.random(h.ne) < p
                S, _ = orc.poll(t, cand, 1)
                G = greedy_matching(cand, t, ends)
                if not np.array_equal(S, G):
                    return f"votes: oracle {np.nonzero(S)[0]} != greedy {np.nonzero(G)[0]} ({n} grains, p={p})"
                for v in range(h.nv):
                    if sum(S[b] for b in vb[v]) > 1:
                        return f"votes: two kept edges at vertex {v}"
                S2, _ = orc.poll(t, S, 1)
                if not np.array_equal(S2, S):
                    return "votes: another round changes the kept edges"
                F, conf = orc.poll(t, cand, 3)
                want = conflict_rule(G, t, ends, vb)
                if not np.array_equal(F, want):
                    return f"conflicts: oracle {np.nonzero(F)[0]} != rule {np.nonzero(want)[0]}"
                for X in range(h.ne):
                    if F[X]:
                        continue
                    a, b = ends[X]
                    if any(F[y] for y in vb[a] if y != X) and any(F[y] for y in vb[b] if y != X):
                        return f"conflicts: non-candidate edge {X} has candidates at both ends"
                F2, _ = orc.poll(t, F, 3)
                if not np.array_equal(F2, F):
                    return "conflicts: the filter is not idempotent"
        orc.close()
    return None


def check_vote_tie(L):
    """Equal t_ext at a vertex: the later boundary slot gets the vote
    (`<=`, topological.h:371), so that edge collects both votes and its
    neighbour is unlabelled."""
    o, h = _mesh(30, 13)
    orc = Orc(L, o, h)
    ends, vb = _topology(h)
    for v in range(0, h.nv, 7):
        for i, j in ((0, 2), (0, 1), (1, 2)):
            a, b = vb[v][i], vb[v][j]
            t = np.full(h.ne, 0.9)
            t[a] = t[b] = 0.5
            cand = np.zeros(h.ne, dtype=bool)
            cand[[a, b]] = True
            S, _ = orc.poll(t, cand, 1)
            if list(np.nonzero(S)[0]) != [b]:
                return f"tie at vertex {v} slots {i},{j}: kept {np.nonzero(S)[0]}, expected [{b}]"
    orc.close()
    return None


def check_conflict_winner(L):
    """Two candidates joined by a non-candidate edge X: the smaller t_ext
    wins; on a tie the smaller id (topological.h:99-111)."""
    o, h = _mesh(30, 13)
    orc = Orc(L, o, h)
    ends, vb = _topology(h)
    cases = 0
    for X in range(0, h.ne, 5):
        p, q = ends[X]
        e1 = next(e for e in vb[p] if e != X)
        e2 = next(e for e in vb[q] if e != X)
        if set(ends[e1]) & set(ends[e2]):
            continue
        for t1, t2 in ((0.3, 0.6), (0.6, 0.3), (0.4, 0.4)):
            t = np.full(h.ne, 0.9)
            t[e1], t[e2] = t1, t2
            cand = np.zeros(h.ne, dtype=bool)
            cand[[e1, e2]] = True
            F, conf = orc.poll(t, cand, 3)
            win = e1 if (t1, e1) < (t2, e2) else e2
            if list(np.nonzero(F)[0]) != [win] or conf[X] != win:
                return f"conflict over edge {X}: kept {np.nonzero(F)[0]}, conf {conf[X]}, expected {win}"
            cases += 1
    orc.close()
    return None if cases > 20 else "too few conflict cases"


# ---------------------------------------------------------------------
# nucleation: closed form, insertion, statistics
# ---------------------------------------------------------------------

def _gb(d_alpha, eps):
    c = math.cos(4 * d_alpha)
    return 1 + 0.5 * eps * (1 - c ** 3)


def _site_geometry(h, A):
    """(D, E, F), (AD, AE, AF), (g1, g2, g3), unit directions and the
    angles theta_1 (F to D), theta_2 (D to E), theta_3 (E to F), counter-
    clockwise, from the mesh alone (diagram nucleation2.cu:83-92)."""
    vb = h["v_bnd"].reshape(-1, 3)
    vg = h["v_grn"].reshape(-1, 3)
    ends = h["b_vtx"].reshape(-1, 2)
    X = h.xy()
    bnd = [int(b) for b in vb[A]]
    nb = [int(ends[b, 1] if ends[b, 0] == A else ends[b, 0]) for b in bnd]
    gs = [set(int(g) for g in vg[v]) for v in (A,) + tuple(nb)]
    g2 = (gs[0] & gs[1] & gs[2]).pop()
    g1 = ((gs[0] & gs[1]) - {g2}).pop()
    g3 = ((gs[0] & gs[2]) - {g2}).pop()
    u, ang = [], []
    for v in nb:
        d = _wrap(X[v] - X[A], h.D)
        u.append(d / np.hypot(*d))
        ang.append(math.atan2(d[1], d[0]))
    tw = 2 * math.pi
    th = [(ang[0] - ang[2]) % tw, (ang[1] - ang[0]) % tw, (ang[2] - ang[1]) % tw]
    return nb, bnd, (g1, g2, g3), u, th


def closed_form_dE(h, o, A, r, ori):
    nb, bnd, (g1, g2, g3), u, th = _site_geometry(h, A)
    se, go = h["g_SE"], h["g_orientation"]
    eps, sc = o.GB_eps, o.GB_scaling
    pair = {bnd[0]: (g1, g2), bnd[1]: (g2, g3), bnd[2]: (g3, g1)}
    gam = [_gb(go[pair[b][0]] - go[pair[b][1]], eps) * sc for b in bnd]
    areas = [0.5 * r * r * math.sin(x) for x in th]
    lens = [2 * r * math.sin(x / 2) for x in th]        # CA, AB, BC
    E0 = r * sum(gam) + areas[0] * se[g1] + areas[1] * se[g2] + areas[2] * se[g3]
    E1 = (_gb(ori - go[g1], eps) * sc * lens[0] + _gb(ori - go[g2], eps) * sc * lens[1] +
          _gb(ori - go[g3], eps) * sc * lens[2])
    return E1 - E0, abs(E1) + abs(E0)


def _nucl_mesh(seed=7, n=40, se_scale=40.0, **kw):
    o, h = _mesh(n, seed, GB_eps=0.3, GB_scaling=1.3, **kw)
    rng = np.random.default_rng(seed)
    h["g_SE"][:] = rng.uniform(0.2, 1.0, h.ng) * se_scale
    h["g_orientation"][:] = rng.uniform(0, 2 * math.pi, h.ng)
    return o, h


def check_nucleation_energy(L):
    o, h = _nucl_mesh()
    orc = Orc(L, o, h)
    X = h.xy()
    rng = np.random.default_rng(5)
    signs = set()
    for A in range(0, h.nv, 3):
        nb, bnd, gs, u, th = _site_geometry(h, A)
        if min(th) <= 0 or abs(sum(th) - 2 * math.pi) > 1e-12:
            return f"site {A}: boundaries not counter-clockwise"
        L0 = min(np.hypot(*_wrap(X[v] - X[A], h.D)) for v in nb)
        for frac in (0.05, 0.3, 0.7):
            r = frac * L0
            ori = rng.uniform(0, 2 * math.pi)
            rc, xy, de, ids = orc.geometry(A, r, ori)
            if rc != 0:
                return f"site {A}: geometry failed ({rc})"
            if list(ids[6:]) != list(gs) or list(ids[:3]) != nb:
                return f"site {A}: grains {list(ids[6:])} / neighbours {list(ids[:3])}, diagram {gs} / {nb}"
            for k in range(3):
                want = np.mod(X[A] + r * u[k], h.D)
                if np.max(np.abs(_wrap(xy[k] - want, h.D))) > 1e-14:
                    return f"site {A}: new vertex {k} at {xy[k]}, expected {want}"
            want, scale = closed_form_dE(h, o, A, r, ori)
            if abs(de - want) > 1e-11 * scale:
                return f"site {A}, r={r}: dE {de} != closed form {want}"
            rc, _, de99, _ = orc.geometry(A, r, ori, do_scale=1)
            want99, scale99 = closed_form_dE(h, o, A, 0.99 * r, ori)
            if abs(de99 - want99) > 1e-11 * scale99:
                return f"site {A}: dE(0.99 r) {de99} != closed form {want99}"
            d2, w2 = de - de99, want - want99
            if abs(w2) > 1e-9 * scale and np.sign(d2) != np.sign(w2):
                return f"site {A}: sign of Delta^2 E {d2} differs from closed form {w2}"
            signs.add(int(np.sign(w2)))
    orc.close()
    return None if signs >= {-1, 1} else f"Delta^2 E signs covered: {signs}"


def _polygon_area(P, D):
    rel = np.cumsum(np.vstack([[0, 0], _wrap(np.diff(P, axis=0), D)]), axis=0)
    x, y = rel[:, 0], rel[:, 1]
    return 0.5 * np.sum(x * np.roll(y, -1) - y * np.roll(x, -1))


def _areas(h):
    gl = h.grain_lists()
    X = h.xy()
    return np.array([_polygon_area(X[gl[g, :h["g_vlen"][g]]], h.D) if h["g_enabled"][g] and h["g_vlen"][g] > 0
                     else 0.0 for g in range(h.ng)])


def _ccw(h, v):
    X = h.xy()
    ends = h["b_vtx"].reshape(-1, 2)
    a = []
    for b in h["v_bnd"].reshape(-1, 3)[v]:
        w = ends[b, 1] if ends[b, 0] == v else ends[b, 0]
        d = _wrap(X[w] - X[v], h.D)
        a.append(math.atan2(d[1], d[0]))
    tw = 2 * math.pi
    return (a[1] - a[0]) % tw + (a[2] - a[1]) % tw + (a[0] - a[2]) % tw < tw + 1e-9


def _cyclic_pair(lst, a, b):
    n = len(lst)
    return any((lst[i] == a and lst[(i + 1) % n] == b) or (lst[i] == b and lst[(i + 1) % n] == a) for i in range(n))


def check_insertion(L):
    o, h = _nucl_mesh(extra_capacity=0.0)
    # room for one nucleus: one grain, three boundaries, two vertices
    big = HostMesh(h.nv + 2, h.ne + 3, h.ng + 1, h.S, h.D)
    for k in big.a:
        src = h.a[k]
        big.a[k][:src.size] = src
    h = big
    X0 = h.xy().copy()
    a0 = _areas(h)
    for A in (0, 11, 23, 40, 57):
        orc = Orc(L, o, h)
        nb, bnd, (g1, g2, g3), u, th = _site_geometry(h, A)
        D_, E_, F_ = nb
        AD, AE, AF = bnd
        L0 = min(np.hypot(*_wrap(X0[v] - X0[A], h.D)) for v in nb)
        r, ori = 0.4 * L0, 1.1
        rc, slots = orc.nucleate(A, r, ori)
        if rc != 0:
            return f"site {A}: nucleate failed ({rc})"
        g, AB, BC, CA, B, Cv = (int(x) for x in slots)
        m = orc.mesh()
        X = m.xy()
        # nucleate builds the nucleus with do_scale = true (nucleation2.cu:526-527):
        # the new vertices sit at 0.99 r
        r = 0.99 * r
        for v, k in ((A, 0), (B, 1), (Cv, 2)):
            want = np.mod(X0[A] + r * u[k], h.D)
            if np.max(np.abs(_wrap(X[v] - want, h.D))) > 1e-14:
                return f"site {A}: vertex {v} at {X[v]}, diagram {want}"
        if not (m["g_enabled"][g] and m["g_type"][g] == 1 and m["g_SE"][g] == 0 and m["g_orientation"][g] == ori):
            return f"site {A}: new grain flags"
        if list(m.grain_lists()[g, :m["g_vlen"][g]]) != [A, B, Cv]:
            return f"site {A}: new grain list {m.grain_lists()[g, :m['g_vlen'][g]]}"
        ends = m["b_vtx"].reshape(-1, 2)
        want_ends = {AB: {A, B}, BC: {B, Cv}, CA: {Cv, A}, AD: {A, D_}, AE: {B, E_}, AF: {Cv, F_}}
        for b, w in want_ends.items():
            if set(ends[b]) != w:
                return f"site {A}: boundary {b} joins {ends[b]}, diagram {w}"
        vbm, vgm = m["v_bnd"].reshape(-1, 3), m["v_grn"].reshape(-1, 3)
        for v, wb, wg in ((A, {AD, AB, CA}, {g1, g2, g}), (B, {AB, BC, AE}, {g, g2, g3}),
                          (Cv, {BC, CA, AF}, {g, g3, g1})):
            if set(vbm[v]) != wb or set(vgm[v]) != wg:
                return f"site {A}: vertex {v} boundaries {vbm[v]} grains {vgm[v]}, diagram {wb} {wg}"
            if not _ccw(m, v):
                return f"site {A}: vertex {v} boundaries not counter-clockwise"
        gl = m.grain_lists()
        lst = {x: list(gl[x, :m["g_vlen"][x]]) for x in (g1, g2, g3)}
        if not (_cyclic_pair(lst[g1], Cv, A) and B not in lst[g1]):
            return f"site {A}: g1 list {lst[g1]}"
        if not (_cyclic_pair(lst[g2], A, B) and Cv not in lst[g2]):
            return f"site {A}: g2 list {lst[g2]}"
        if not (_cyclic_pair(lst[g3], B, Cv) and A not in lst[g3]):
            return f"site {A}: g3 list {lst[g3]}"
        a1 = _areas(m)
        small = [0.5 * r * r * math.sin(x) for x in th]
        for gg, d in ((g1, small[0]), (g2, small[1]), (g3, small[2])):
            if abs((abs(a0[gg]) - abs(a1[gg])) - d) > 1e-12:
                return f"site {A}: grain {gg} lost {abs(a0[gg]) - abs(a1[gg])}, diagram {d}"
        if abs(abs(a1[g]) - sum(small)) > 1e-12:
            return f"site {A}: new grain area {a1[g]}, diagram {sum(small)}"
        ta = C.c_double()
        bad = L.orc_check_mesh(C.byref(m.struct), C.byref(ta))
        euler = int(m["v_enabled"].sum()) - int(m["b_enabled"].sum()) + int(m["g_enabled"].sum())
        if bad != 0 or euler != 0 or abs(ta.value - h.D ** 2) > 1e-12:
            return f"site {A}: check_data {bad}, Euler {euler}, area {ta.value}"
        orc.close()
    return None


def check_metropolis(L, n_trials=3000):
    o, h = _nucl_mesh(se_scale=12.0, nucleation_mode=sevc.NUCL_MC, MC_k=20.0, do_nucleation=1, seed=77)
    h["g_SE"][:] = 12.0
    orc = Orc(L, o, h)
    sites = np.zeros(h.nv, dtype=np.int64)
    pm, am = [], []
    for n in range(n_trials):
        acc, site, de, nf = orc.mc_trial(n)
        sites[site] += 1
        if de < 0:
            if acc != 1:
                return f"step {n}: dE < 0 not accepted"
        elif nf < 0:
            pm.append(min(1.0, math.exp(-o.MC_k * de)))
            am.append(acc)
        elif acc != 0:
            return f"step {n}: accepted with dE >= 0 and Delta^2 E >= 0"
    orc.close()
    pm, am = np.array(pm), np.array(am)
    if pm.size < 300 or not (0.05 < pm.mean() < 0.95):
        return f"Metropolis branch poorly covered: {pm.size} trials, mean p {pm.mean() if pm.size else 0}"
    z = (am.sum() - pm.sum()) / math.sqrt(np.sum(pm * (1 - pm)))
    if abs(z) > 4.5:
        return f"Metropolis acceptance count off by z = {z:.2f}"
    order = np.argsort(pm)
    chi = 0.0
    for part in np.array_split(order, 6):
        e, var = pm[part].sum(), np.sum(pm[part] * (1 - pm[part]))
        chi += (am[part].sum() - e) ** 2 / var
    if stats.chi2.sf(chi, 6) < 1e-4:
        return f"Metropolis acceptance by probability bin: chi2 {chi:.1f}"
    if stats.chisquare(sites).pvalue < 1e-4:
        return "random site not uniform over alive vertices"
    return None


def check_rate(L, n_steps=2500):
    o, h = _nucl_mesh(se_scale=1.0, nucleation_mode=sevc.NUCL_RATE, nucleation_rate=0.02, do_nucleation=1,
                      seed=91, nucleation_scaling=0.9)
    h["g_SE"][:] = np.random.default_rng(3).uniform(0.0, 20.0, h.ng)
    orc = Orc(L, o, h)
    p = 0.02 * h["g_SE"]
    cnt = np.zeros(h.ng, dtype=np.int64)
    slot = np.zeros(h.ng * h.S, dtype=np.int64)
    gl = h.grain_lists()
    for n in range(n_steps):
        gs, vs = orc.rate_draws(n, h.ng)
        for g, v in zip(gs, vs):
            cnt[g] += 1
            k = np.nonzero(gl[g, :h["g_vlen"][g]] == v)[0]
            if k.size != 1:
                return f"step {n}: site {v} is not a vertex of grain {g}"
            slot[g * h.S + k[0]] += 1
    orc.close()
    z = (cnt.sum() - n_steps * p.sum()) / math.sqrt(n_steps * np.sum(p * (1 - p)))
    if abs(z) > 4.5:
        return f"rate events {cnt.sum()} vs expected {n_steps * p.sum():.0f} (z = {z:.2f})"
    chi = np.sum((cnt - n_steps * p) ** 2 / (n_steps * p * (1 - p)))
    if stats.chi2.sf(chi, h.ng) < 1e-4:
        return f"per-grain event counts: chi2 {chi:.1f} over {h.ng} grains"
    chi, dof = 0.0, 0
    for g in range(h.ng):
        n = h["g_vlen"][g]
        c = slot[g * h.S:g * h.S + n]
        if c.sum() < 5 * n:
            continue
        e = c.sum() / n
        chi += np.sum((c - e) ** 2 / e)
        dof += n - 1
    if dof == 0 or stats.chi2.sf(chi, dof) < 1e-4:
        return f"site within a grain not uniform: chi2 {chi:.1f}, dof {dof}"
    return None


CHECKS = [check_polling, check_vote_tie, check_conflict_winner, check_nucleation_energy, check_insertion,
          check_metropolis, check_rate]


@pytest.fixture(scope="module")
def oracle():
    oracle_lib()        # builds the oracle if needed
    return _bind(C.CDLL(ORACLE_SO))


@pytest.mark.parametrize("check", CHECKS, ids=[c.__name__ for c in CHECKS])
def test_oracle_pin(oracle, check):
    assert check(oracle) is None


# Planted mistakes (ORC_MUTANT=k in oracle/sevc_oracle.c) and what they break.
MUTANTS = {
    1: "vote tie-break `<` for `<=` (topological.h:371)",
    2: "vote rounds stop after one round",
    3: "conflict won by the larger t_ext",
    4: "conflict tie won by the larger id",
    5: "stored-energy terms of g1 and g2 swapped in E0",
    6: "sign of the area terms of E0",
    7: "Metropolis test reversed (u > exp(-k dE))",
    8: "g1 instead of g3 replaced by the new grain at the site",
    9: "new vertices B and C placed on AF and AE",
    10: "rate-mode probability without the stored energy",
    11: "lengths CA and AB attributed to g2 and g1 in E1",
}


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")
    src = [os.path.join(REPO, "oracle", f) for f in ("sevc_oracle.c", "sevc_oracle_io.c")]
    procs = {}
    for k in MUTANTS:
        out = str(d / f"liboracle_mut{k}.so")
        procs[k] = subprocess.Popen(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
                                     f"-DORC_MUTANT={k}", "-shared", "-o", out] + src + ["-lm"])
    for k, p in procs.items():
        assert p.wait() == 0, k
    return d


@pytest.mark.parametrize("k", sorted(MUTANTS), ids=[f"mutant{k}" for k in sorted(MUTANTS)])
def test_planted_mistake_fails_a_pin(mutant_dir, k):
    L = _bind(C.CDLL(str(mutant_dir / f"liboracle_mut{k}.so")))
    failures = []
    for check in CHECKS:
        msg = check(L)
        if msg:
            failures.append(f"{check.__name__}: {msg}")
            break
    assert failures, f"mutant {k} ({MUTANTS[k]}) passes every pin"
    print(f"mutant {k} ({MUTANTS[k]}) caught by {failures[0]}")
