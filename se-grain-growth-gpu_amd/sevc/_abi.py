"""ctypes mirror of include/sevc.h (structs, constants) and numpy-backed
host meshes.  Shared by the product binding (sevc.lib) and by the tests,
which hand the same structs to the oracle library."""
import ctypes as C

import numpy as np

OK = 0
E_INVALID_ARG = -1
E_IO = -2
E_TOPOLOGY = -3
E_HIP = -4
E_NOMEM = -5
E_GRAIN_FULL = -6
E_STATE = -7
E_NO_DEVICE = -8
E_CHECK = -9

NUCL_MC = 0
NUCL_RATE = 1
VEL_FD = 0
VEL_ANALYTIC = 1
FNAME = 256
REF_MAX_VRT_PER_GRN = 100


class Options(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("dtau", C.c_double), ("nucleation_gap", C.c_int32),
        ("tmax", C.c_double), ("nmax", C.c_int32), ("snap_gap", C.c_int32),
        ("leave_percent", C.c_double), ("n_grains", C.c_int32),
        ("vertex_fname", C.c_char * FNAME), ("boundary_fname", C.c_char * FNAME),
        ("orientation_fname", C.c_char * FNAME), ("SE_fname", C.c_char * FNAME),
        ("GB_eps", C.c_double), ("GB_scaling", C.c_double), ("SE_eps", C.c_double),
        ("SE_min", C.c_double), ("SE_max", C.c_double), ("nucleation_scaling", C.c_double),
        ("domain_bound", C.c_double), ("do_nucleation", C.c_int32), ("MC_k", C.c_double),
        ("n_blocks", C.c_int32), ("n_threads", C.c_int32), ("debug_lvl", C.c_int32),
        ("output_folder", C.c_char * FNAME),
        ("mobility", C.c_double), ("adaptive_dt", C.c_int32), ("dt_safety", C.c_double),
        ("dt_min", C.c_double), ("flip_threshold", C.c_double), ("nucleation_mode", C.c_int32),
        ("nucleation_rate", C.c_double), ("seed", C.c_uint64), ("max_vrt_per_grn", C.c_int32),
        ("extra_capacity", C.c_double), ("velocity_model", C.c_int32), ("events", C.c_int32),
        ("nucleation_gate", C.c_int32),
    ]

    def copy(self):
        o = Options()
        C.memmove(C.byref(o), C.byref(self), C.sizeof(Options))
        return o


class Mesh(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int32), ("n_boundaries", C.c_int32), ("n_grains", C.c_int32),
        ("max_vrt_per_grn", C.c_int32), ("domain_bound", C.c_double),
        ("pos", C.POINTER(C.c_double)), ("vel", C.POINTER(C.c_double)),
        ("v_bnd", C.POINTER(C.c_int32)), ("v_grn", C.POINTER(C.c_int32)),
        ("v_enabled", C.POINTER(C.c_uint8)), ("v_energy", C.POINTER(C.c_double)),
        ("v_vel_parts", C.POINTER(C.c_double)),
        ("b_vtx", C.POINTER(C.c_int32)), ("b_energy", C.POINTER(C.c_double)),
        ("b_arclength", C.POINTER(C.c_double)), ("b_t_ext", C.POINTER(C.c_double)),
        ("b_enabled", C.POINTER(C.c_uint8)), ("b_candidate", C.POINTER(C.c_uint8)),
        ("g_vtx", C.POINTER(C.c_int32)), ("g_vlen", C.POINTER(C.c_int32)),
        ("g_area", C.POINTER(C.c_double)), ("g_dAdt", C.POINTER(C.c_double)),
        ("g_orientation", C.POINTER(C.c_double)), ("g_SE", C.POINTER(C.c_double)),
        ("g_enabled", C.POINTER(C.c_uint8)), ("g_type", C.POINTER(C.c_uint8)),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("steps_done", C.c_int32), ("n", C.c_int32), ("t", C.c_double), ("dt_last", C.c_double),
        ("n_grains", C.c_int32), ("n_vertices", C.c_int32), ("nucleated_area_pct", C.c_double),
        ("done", C.c_int32), ("status", C.c_int32),
        ("n_flips", C.c_int64), ("n_removals", C.c_int64), ("n_two_grain", C.c_int64),
        ("n_nucleations", C.c_int64), ("n_vertex_steps", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Grads(C.Structure):
    _fields_ = [("d_pos", C.POINTER(C.c_double)), ("d_SE", C.POINTER(C.c_double)),
                ("d_gb_scaling", C.c_double), ("d_mobility", C.c_double)]


# (field, dtype, per-entity multiplicity, entity) of every Mesh array
MESH_ARRAYS = [
    ("pos", np.float64, 2, "v"), ("vel", np.float64, 2, "v"),
    ("v_bnd", np.int32, 3, "v"), ("v_grn", np.int32, 3, "v"), ("v_enabled", np.uint8, 1, "v"),
    ("v_energy", np.float64, 1, "v"), ("v_vel_parts", np.float64, 4, "v"),
    ("b_vtx", np.int32, 2, "b"), ("b_energy", np.float64, 1, "b"), ("b_arclength", np.float64, 1, "b"),
    ("b_t_ext", np.float64, 1, "b"), ("b_enabled", np.uint8, 1, "b"), ("b_candidate", np.uint8, 1, "b"),
    ("g_vtx", np.int32, "S", "g"), ("g_vlen", np.int32, 1, "g"), ("g_area", np.float64, 1, "g"),
    ("g_dAdt", np.float64, 1, "g"), ("g_orientation", np.float64, 1, "g"), ("g_SE", np.float64, 1, "g"),
    ("g_enabled", np.uint8, 1, "g"), ("g_type", np.uint8, 1, "g"),
]

_CT = {np.float64: C.c_double, np.int32: C.c_int32, np.uint8: C.c_uint8}


class HostMesh:
    """A Mesh whose arrays are numpy arrays owned by Python."""

    def __init__(self, nv, ne, ng, S, D=1.0):
        self.nv, self.ne, self.ng, self.S, self.D = int(nv), int(ne), int(ng), int(S), float(D)
        self.a = {}
        for name, dt, mult, ent in MESH_ARRAYS:
            n = {"v": self.nv, "b": self.ne, "g": self.ng}[ent]
            m = self.S if mult == "S" else mult
            arr = np.zeros(n * m, dtype=dt)
            if dt is np.int32:
                arr[:] = -1
            self.a[name] = arr
        self.a["b_t_ext"][:] = 100000.0
        self._struct = None

    @property
    def struct(self):
        m = Mesh()
        m.n_vertices, m.n_boundaries, m.n_grains = self.nv, self.ne, self.ng
        m.max_vrt_per_grn, m.domain_bound = self.S, self.D
        for name, dt, _, _ in MESH_ARRAYS:
            setattr(m, name, self.a[name].ctypes.data_as(C.POINTER(_CT[dt])))
        self._struct = m
        return m

    def __getitem__(self, k):
        return self.a[k]

    def copy(self):
        h = HostMesh(self.nv, self.ne, self.ng, self.S, self.D)
        for k, v in self.a.items():
            h.a[k] = v.copy()
        return h

    @classmethod
    def from_struct(cls, m):
        """Copy a Mesh filled by C code into numpy arrays."""
        h = cls(m.n_vertices, m.n_boundaries, m.n_grains, m.max_vrt_per_grn, m.domain_bound)
        for name, dt, _, _ in MESH_ARRAYS:
            p = getattr(m, name)
            if p:
                n = h.a[name].size
                h.a[name][:] = np.ctypeslib.as_array(p, shape=(n,))
        return h

    # convenient views
    def xy(self):
        return self.a["pos"].reshape(-1, 2)

    def grain_lists(self):
        return self.a["g_vtx"].reshape(self.ng, self.S)


def default_options():
    """Python restatement of sevc_options_default (options.h:71-85 plus the
    extension defaults).  The C function is authoritative; a test checks
    that the two agree."""
    o = Options()
    o.dt = 1e-4
    o.dtau = 1e-4
    o.nucleation_gap = 1
    o.tmax = -1.0
    o.nmax = -1
    o.snap_gap = 100
    o.leave_percent = -1.0
    o.n_grains = 0
    o.GB_eps = 0.02
    o.GB_scaling = 1.0
    o.SE_eps = 1.0
    o.SE_min = 999.0
    o.SE_max = 999.0
    o.nucleation_scaling = 0.9
    o.domain_bound = 1.0
    o.do_nucleation = 1
    o.MC_k = 1.0
    o.n_blocks = 32
    o.n_threads = 256
    o.debug_lvl = 0
    o.mobility = 1.0
    o.adaptive_dt = 0
    o.dt_safety = 0.5
    o.dt_min = 0.0
    o.flip_threshold = 0.0
    o.nucleation_mode = NUCL_MC
    o.nucleation_rate = 0.0
    o.seed = 32
    o.max_vrt_per_grn = 32
    o.extra_capacity = 0.0
    o.velocity_model = VEL_FD
    o.events = 1
    o.nucleation_gate = -1
    return o
