"""Test helpers: the oracle binding (test infrastructure) and state
comparison utilities."""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
REF2K = os.path.join(GOLDEN, "ref2k")
ORACLE_SO = os.path.join(REPO, "oracle", "build", "liboracle.so")
ORACLE_OMP_SO = os.path.join(REPO, "oracle", "build", "liboracle_omp.so")

import sevc  # noqa: E402
from sevc._abi import HostMesh, Mesh, Options, StepStats  # noqa: E402

_orc = {}


def oracle_lib(omp=False):
    """The serial oracle, or (omp=True) its OpenMP build, a timing baseline
    that gives the same results."""
    path = ORACLE_OMP_SO if omp else ORACLE_SO
    if path not in _orc:
        if not os.path.exists(path):
            subprocess.run(["make", "-C", os.path.join(REPO, "oracle")], check=True, capture_output=True)
        L = C.CDLL(path)
        P = C.POINTER
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [P(Options), P(Mesh)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p, C.c_int32, P(StepStats)]
        L.orc_get_mesh.argtypes = [C.c_void_p, P(Mesh)]
        L.orc_compute_velocities.argtypes = [C.c_void_p]
        L.orc_request_diagnostics.argtypes = [C.c_void_p]
        L.orc_load_files.argtypes = [P(Options), P(Mesh)]
        L.orc_build_mesh.argtypes = [P(Options), C.c_int32, P(C.c_double), C.c_int32, P(C.c_int32), C.c_int32,
                                     P(C.c_double), P(C.c_double), P(Mesh)]
        L.orc_mesh_free.argtypes = [P(Mesh)]
        L.orc_check_mesh.argtypes = [P(Mesh), P(C.c_double)]
        L.orc_wrap_dist.restype = C.c_double
        L.orc_wrap_dist.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_dom_mod.restype = C.c_double
        L.orc_dom_mod.argtypes = [C.c_double, C.c_double]
        L.orc_mag.restype = C.c_double
        L.orc_mag.argtypes = [C.c_double, C.c_double]
        L.orc_polygon_area.restype = C.c_double
        L.orc_polygon_area.argtypes = [P(C.c_double), P(C.c_double), C.c_int32, C.c_double]
        L.orc_sum_chunked.restype = C.c_double
        L.orc_sum_chunked.argtypes = [P(C.c_double), P(C.c_uint8), C.c_int32]
        L.orc_atan2.restype = C.c_double
        L.orc_atan2.argtypes = [C.c_double, C.c_double]
        for f in ("orc_sin", "orc_cos", "orc_exp"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_double]
        L.orc_philox.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        L.orc_window_forward.argtypes = [C.c_void_p, C.c_int32]
        L.orc_adjoint_window.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_double),
                                         P(C.c_double), P(C.c_double)]
        _orc[path] = L
    return _orc[path]


class Oracle:
    """The serial FP64 restatement of the reference step."""

    def __init__(self, opts, mesh, omp=False):
        self.L = oracle_lib(omp)
        self.template = mesh
        self.h = self.L.orc_create(C.byref(opts), C.byref(mesh.struct))
        if not self.h:
            raise RuntimeError("orc_create failed")

    def step(self, n):
        st = StepStats()
        rc = self.L.orc_step(self.h, int(n), C.byref(st))
        return rc, st

    def velocities(self):
        return self.L.orc_compute_velocities(self.h)

    def request_diagnostics(self):
        self.L.orc_request_diagnostics(self.h)

    def window_forward(self, steps):
        return self.L.orc_window_forward(self.h, int(steps))

    def adjoint(self, dL_dpos):
        """Reverse pass over the last window: (d_pos, d_SE, d_gb_scaling, d_mobility)."""
        t = self.template
        P = C.POINTER(C.c_double)
        g = np.ascontiguousarray(dL_dpos, dtype=np.float64).reshape(-1)
        d_pos = np.zeros(2 * t.nv)
        d_se = np.zeros(t.ng)
        gb, mu = C.c_double(), C.c_double()
        rc = self.L.orc_adjoint_window(self.h, g.ctypes.data_as(P), d_pos.ctypes.data_as(P),
                                       d_se.ctypes.data_as(P), C.byref(gb), C.byref(mu))
        assert rc == 0, rc
        return d_pos, d_se, gb.value, mu.value

    def get_mesh(self):
        t = self.template
        out = HostMesh(t.nv, t.ne, t.ng, t.S, t.D)
        rc = self.L.orc_get_mesh(self.h, C.byref(out.struct))
        assert rc == 0
        return out

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def oracle_load(opts):
    m = Mesh()
    rc = oracle_lib().orc_load_files(C.byref(opts), C.byref(m))
    if rc != 0:
        return rc, None
    h = HostMesh.from_struct(m)
    oracle_lib().orc_mesh_free(C.byref(m))
    return 0, h


def ref2k_options(se="uniform", **kw):
    o = sevc.default_options()
    o.vertex_fname = os.path.join(REF2K, "vertices.txt").encode()
    o.boundary_fname = os.path.join(REF2K, "ridges.txt").encode()
    o.orientation_fname = os.path.join(REF2K, "oris.txt").encode()
    o.SE_fname = os.path.join(REF2K, "SE", se + ".txt").encode()
    o.n_grains = 2000
    o.GB_eps = 0.0
    o.SE_eps = 0.5
    o.do_nucleation = 0
    o.dt = 1e-5
    for k, v in kw.items():
        setattr(o, k, v)
    return o


# Fields compared bit for bit after a step.  Diagnostic-only arrays are
# compared where the test asks for them.
STATE_FIELDS = ["pos", "vel", "v_bnd", "v_grn", "v_enabled", "b_vtx", "b_enabled", "b_candidate", "b_t_ext",
                "b_arclength", "b_energy", "g_vtx", "g_vlen", "g_enabled", "g_area", "g_SE", "g_orientation",
                "g_type"]


def diff_fields(a, b, fields=STATE_FIELDS):
    bad = {}
    for f in fields:
        x, y = a.a[f], b.a[f]
        if x.dtype.kind == "f":
            neq = ~((x == y) | (np.isnan(x) & np.isnan(y)))
            # +0.0 and -0.0 compare equal; require identical bits as well
            neq |= (x.view(np.uint64) != y.view(np.uint64))
        else:
            neq = x != y
        if neq.any():
            bad[f] = int(neq.sum())
    return bad


def oracle_check(m):
    """The oracle's check_data count of a host mesh."""
    ta = C.c_double()
    return oracle_lib().orc_check_mesh(C.byref(m.struct), C.byref(ta))


def planted_corruptions(m):
    """Corrupted copies of a valid mesh whose slots are all enabled, each
    with its violation count worked out by hand from check_data's
    conditions (utils.h:85-336) and the out-of-range rule (an id outside
    [0, n) counts once for the entity holding it, and the checks that would
    follow it are skipped).  Returns (name, mesh, expected) triples."""
    out = []
    nv, ne = m.nv, m.ne
    S = m.S

    # edge 10 disabled: both ends point to a disabled boundary (utils.h:190)
    # and each of the 4 other edges at those ends has a disabled neighbour
    # edge on that side (:132-145)
    c = m.copy()
    c["b_enabled"][10] = 0
    out.append(("edge disabled", c, 6))

    # vertex 20 disabled: its 3 edges have a disabled end (:99), its 3
    # grains list a disabled vertex (:280)
    c = m.copy()
    c["v_enabled"][20] = 0
    out.append(("vertex disabled", c, 6))

    # grain 30 with vlen one short: count mismatch (:303-316), and the last
    # vertex no longer finds the grain's list holding it (:236-253)
    c = m.copy()
    c["g_vlen"][30] -= 1
    out.append(("grain list one short", c, 2))

    # vertex 40's first boundary id out of range: the vertex (1); that
    # boundary's end no longer lists it (:107-130) and has an out-of-range
    # neighbour (:132-145) (2); the vertex's two other edges see it (2)
    c = m.copy()
    c["v_bnd"][3 * 40] = ne + 5
    out.append(("boundary id out of range at a vertex", c, 5))

    # grain 50's first list entry out of range: the grain (:280), and the
    # vertex that was there no longer finds itself in the grain (:236-253)
    c = m.copy()
    c["g_vtx"][50 * S] = nv + 3
    out.append(("vertex id out of range in a grain list", c, 2))

    # edge 60's first end out of range: the edge (1), and its old end is no
    # longer referenced by one of its boundaries (:227-234)
    c = m.copy()
    c["b_vtx"][2 * 60] = nv + 1
    out.append(("edge end out of range", c, 2))
    return out


def random_corruptions(m, n, seed):
    """n copies of m, each with one to three random entries of the
    topology arrays overwritten by a random id (in range or not) or flag."""
    rng = np.random.default_rng(seed)
    fields = ["v_bnd", "v_grn", "v_enabled", "b_vtx", "b_enabled", "g_vtx", "g_vlen", "g_enabled"]
    hi = {"v_bnd": m.ne, "v_grn": m.ng, "b_vtx": m.nv, "g_vtx": m.nv, "g_vlen": m.S}
    out = []
    for _ in range(n):
        c = m.copy()
        for _ in range(int(rng.integers(1, 4))):
            f = fields[int(rng.integers(len(fields)))]
            k = int(rng.integers(c[f].size))
            if f in hi:
                c[f][k] = int(rng.integers(-2, hi[f] + 3))
            else:
                c[f][k] = 1 - c[f][k]
        out.append(c)
    return out
