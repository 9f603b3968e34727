"""ctypes binding of libsevc.so (the C ABI of include/sevc.h).

The product path is the native library: importing this module fails loudly
when it has not been built, and nothing here falls back to Python or CPU
arithmetic."""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import HostMesh, Mesh, Options, StepStats

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SEVC_LIB_PATH") or os.path.join(os.path.dirname(_HERE), "build", "libsevc.so")
CLI_PATH = os.path.join(os.path.dirname(LIB_PATH), "sevc")       # command-line driver (sevc -c config.ini)

_lib = None


class SevcError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"sevc error {code}: {msg}")
        self.code = code


def lib():
    """Load libsevc.so (built in-tree by sevc.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    sig = {
        "sevc_options_default": (None, [P(Options)]),
        "sevc_options_from_ini": (C.c_int, [C.c_char_p, P(Options)]),
        "sevc_mesh_alloc": (C.c_int, [P(Mesh), C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
        "sevc_mesh_free": (None, [P(Mesh)]),
        "sevc_mesh_load_files": (C.c_int, [P(Options), P(Mesh)]),
        "sevc_mesh_build": (C.c_int, [P(Options), C.c_int32, P(C.c_double), C.c_int32, P(C.c_int32),
                                      C.c_int32, P(C.c_double), P(C.c_double), P(Mesh)]),
        "sevc_voronoi_generate": (C.c_int, [C.c_int32, C.c_uint64, P(C.c_int32), P(P(C.c_double)),
                                            P(C.c_int32), P(P(C.c_int32))]),
        "sevc_free": (None, [C.c_void_p]),
        "sevc_check_mesh": (C.c_int, [P(Mesh), P(C.c_double)]),
        "sevc_check_mesh_device": (C.c_int, [P(Mesh), P(C.c_int32)]),
        "sevc_check_state": (C.c_int, [C.c_void_p, P(C.c_int32)]),
        "sevc_check_meshes_device_noalloc": (C.c_int, [P(Mesh), C.c_int32, C.c_void_p, C.c_void_p]),
        "sevc_export_snapshot": (C.c_int, [P(Mesh), P(Options), C.c_char_p, C.c_int32, C.c_double]),
        "sevc_load_snapshot": (C.c_int, [C.c_char_p, C.c_int32, P(Mesh), P(C.c_double)]),
        "sevc_create": (C.c_int, [P(Options), P(Mesh), P(C.c_void_p)]),
        "sevc_batch_create": (C.c_int, [P(Options), P(Mesh), C.c_int32, P(C.c_void_p)]),
        "sevc_step": (C.c_int, [C.c_void_p, C.c_int32, P(StepStats)]),
        "sevc_step_graph": (C.c_int, [C.c_void_p, C.c_int32, P(StepStats)]),
        "sevc_member_stats": (C.c_int, [C.c_void_p, C.c_int32, P(StepStats)]),
        "sevc_member_sizes": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_int32),
                                        P(C.c_int32)]),
        "sevc_n_members": (C.c_int32, [C.c_void_p]),
        "sevc_get_mesh": (C.c_int, [C.c_void_p, C.c_int32, P(Mesh)]),
        "sevc_set_member_params": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_double), P(C.c_double),
                                             C.c_double, C.c_double]),
        "sevc_request_diagnostics": (C.c_int, [C.c_void_p]),
        "sevc_set_member_mesh": (C.c_int, [C.c_void_p, C.c_int32, P(Mesh)]),
        "sevc_checkpoint": (C.c_int, [C.c_void_p, C.c_int32]),
        "sevc_restore": (C.c_int, [C.c_void_p, C.c_int32]),
        "sevc_get_positions": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_double)]),
        "sevc_autotune": (C.c_int, [C.c_void_p, P(C.c_int32)]),
        "sevc_velocity_config": (C.c_int, [C.c_void_p, P(C.c_int32), P(C.c_int32)]),
        "sevc_set_velocity_config": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
        "sevc_set_poll_limit": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
        "sevc_set_search_round": (C.c_int, [C.c_void_p, C.c_int32]),
        "sevc_set_timing_sample": (C.c_int, [C.c_void_p, C.c_int32]),
        "sevc_area_histogram": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_double), P(C.c_double), P(C.c_double),
                                          P(C.c_double), P(P(C.c_double)), C.c_int32]),
        "sevc_window_forward": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
        "sevc_adjoint_window": (C.c_int, [C.c_void_p, P(P(C.c_double)), P(_abi.Grads)]),
        "sevc_enable_kernel_timing": (C.c_int, [C.c_void_p, C.c_int32]),
        "sevc_kernel_times": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, P(C.c_double), P(C.c_int64)]),
        "sevc_destroy": (None, [C.c_void_p]),
        "sevc_last_error": (C.c_char_p, [C.c_void_p]),
        "sevc_strerror": (C.c_char_p, [C.c_int]),
        "sevc_device_count": (C.c_int, []),
        "sevc_selftest_math": (C.c_int, [C.c_int32, C.c_int32, P(C.c_double), P(C.c_double), P(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
    _lib = L
    return L


def _check(rc, state=None):
    if rc != 0:
        msg = lib().sevc_strerror(rc).decode()
        if state:
            msg += " / " + lib().sevc_last_error(state).decode()
        raise SevcError(rc, msg)


def default_options():
    o = Options()
    lib().sevc_options_default(C.byref(o))
    return o


def options_from_ini(path):
    """Options of an .ini file: the defaults, then every key the file sets."""
    o = Options()
    _check(lib().sevc_options_from_ini(path.encode(), C.byref(o)))
    return o


def _take_mesh(m):
    h = HostMesh.from_struct(m)
    lib().sevc_mesh_free(C.byref(m))
    return h


def load_files(opts):
    m = Mesh()
    _check(lib().sevc_mesh_load_files(C.byref(opts), C.byref(m)))
    return _take_mesh(m)


def build_mesh(opts, unit_pos, ridges, orientations01, se_raw, n_grains_expected=0):
    unit_pos = np.ascontiguousarray(unit_pos, dtype=np.float64).reshape(-1)
    ridges = np.ascontiguousarray(ridges, dtype=np.int32).reshape(-1)
    ori = np.ascontiguousarray(orientations01, dtype=np.float64)
    se = np.ascontiguousarray(se_raw, dtype=np.float64)
    m = Mesh()
    P = C.POINTER
    _check(lib().sevc_mesh_build(C.byref(opts), unit_pos.size // 2, unit_pos.ctypes.data_as(P(C.c_double)),
                                 ridges.size // 2, ridges.ctypes.data_as(P(C.c_int32)), int(n_grains_expected),
                                 ori.ctypes.data_as(P(C.c_double)), se.ctypes.data_as(P(C.c_double)),
                                 C.byref(m)))
    return _take_mesh(m)


def load_snapshot(path, max_vrt_per_grn=32):
    """(mesh, time) of a snapshot written by sevc_export_snapshot."""
    m = Mesh()
    t = C.c_double()
    _check(lib().sevc_load_snapshot(str(path).encode(), int(max_vrt_per_grn), C.byref(m), C.byref(t)))
    return _take_mesh(m), t.value


def voronoi(n_seeds, seed):
    """Periodic Voronoi tessellation: (unit vertex positions (nv,2), ridges (nr,2))."""
    nv, nr = C.c_int32(), C.c_int32()
    pp, rp = C.POINTER(C.c_double)(), C.POINTER(C.c_int32)()
    _check(lib().sevc_voronoi_generate(int(n_seeds), int(seed), C.byref(nv), C.byref(pp), C.byref(nr),
                                       C.byref(rp)))
    pos = np.ctypeslib.as_array(pp, shape=(nv.value * 2,)).copy().reshape(-1, 2)
    rid = np.ctypeslib.as_array(rp, shape=(nr.value * 2,)).copy().reshape(-1, 2)
    lib().sevc_free(C.cast(pp, C.c_void_p))
    lib().sevc_free(C.cast(rp, C.c_void_p))
    return pos, rid


def check_mesh(h):
    ta = C.c_double()
    n = lib().sevc_check_mesh(C.byref(h.struct), C.byref(ta))
    return n, ta.value


def check_mesh_device(h):
    """check_data's violation count of a host mesh, counted on the GPU."""
    bad = C.c_int32()
    _check(lib().sevc_check_mesh_device(C.byref(h.struct), C.byref(bad)))
    return bad.value


class State:
    """A device state holding one or more members (sevc_batch_create)."""

    def __init__(self, opts, meshes):
        if isinstance(meshes, HostMesh):
            meshes, opts = [meshes], [opts]
        self.meshes_in = meshes
        arr_o = (Options * len(meshes))(*[o.copy() for o in opts])
        self._structs = [m.struct for m in meshes]
        arr_m = (Mesh * len(meshes))(*self._structs)
        h = C.c_void_p()
        rc = lib().sevc_batch_create(arr_o, arr_m, len(meshes), C.byref(h))
        if rc != 0:
            raise SevcError(rc, lib().sevc_strerror(rc).decode() + " / " + lib().sevc_last_error(None).decode())
        self.h = h
        self.n_members = len(meshes)

    def step(self, n, graph=False):
        st = StepStats()
        f = lib().sevc_step_graph if graph else lib().sevc_step
        _check(f(self.h, int(n), C.byref(st)), self.h)
        return st

    def check(self):
        """check_data's violation count of every member, counted on the
        device state in place."""
        bad = (C.c_int32 * self.n_members)()
        _check(lib().sevc_check_state(self.h, bad), self.h)
        return list(bad)

    def member_stats(self, m):
        st = StepStats()
        _check(lib().sevc_member_stats(self.h, int(m), C.byref(st)), self.h)
        return st

    def sizes(self, m=0):
        a = [C.c_int32() for _ in range(4)]
        _check(lib().sevc_member_sizes(self.h, int(m), *[C.byref(x) for x in a]), self.h)
        return tuple(x.value for x in a)

    def get_mesh(self, m=0):
        nv, ne, ng, S = self.sizes(m)
        h = HostMesh(nv, ne, ng, S)
        _check(lib().sevc_get_mesh(self.h, int(m), C.byref(h.struct)), self.h)
        h.D = h.struct.domain_bound
        return h

    def area_histogram(self, n_bins, ranges, targets, seed=False, want_grad=False):
        """Device area-histogram loss of every member at the current
        positions.  ranges: (n_members, 3) of (a_min, a_max, width);
        targets: (n_members, n_bins).  Returns (hist, loss[, dL_dpos list]);
        seed=True keeps dL/dx on the device for adjoint(None)."""
        P = C.POINTER(C.c_double)
        B = self.n_members
        rg = np.ascontiguousarray(ranges, dtype=np.float64).reshape(B, 3)
        tg = np.ascontiguousarray(targets, dtype=np.float64).reshape(B, n_bins)
        hist = np.zeros((B, n_bins))
        loss = np.zeros(B)
        grads, arr = None, None
        if want_grad:
            grads = [np.zeros(2 * self.sizes(m)[0]) for m in range(B)]
            arr = (P * B)(*[g.ctypes.data_as(P) for g in grads])
        _check(lib().sevc_area_histogram(self.h, int(n_bins), rg.ctypes.data_as(P), tg.ctypes.data_as(P),
                                         hist.ctypes.data_as(P), loss.ctypes.data_as(P), arr, int(bool(seed))),
               self.h)
        return (hist, loss, grads) if want_grad else (hist, loss)

    def autotune(self):
        """Time the velocity grain pass for several tile sizes and keep the
        fastest; returns the chosen grains per tile."""
        t = C.c_int32()
        _check(lib().sevc_autotune(self.h, C.byref(t)), self.h)
        return t.value

    def velocity_config(self):
        """(path, tile): 0 = two passes (grains per tile), 1 = one pass over
        vertex tiles (vertices per tile)."""
        p, t = C.c_int32(), C.c_int32()
        _check(lib().sevc_velocity_config(self.h, C.byref(p), C.byref(t)), self.h)
        return p.value, t.value

    def set_poll_limit(self, max_fast, max_lds=16384):
        """Candidate lists up to max_fast poll from registers, up to max_lds
        from LDS, longer ones with the general rounds (same results)."""
        _check(lib().sevc_set_poll_limit(self.h, int(max_fast), int(max_lds)), self.h)

    def set_search_round(self, blocks):
        """Blocks of each kind read per round of the nucleation free-slot
        search (0 = default; same results)."""
        _check(lib().sevc_set_search_round(self.h, int(blocks)), self.h)

    def set_velocity_config(self, path, tile):
        _check(lib().sevc_set_velocity_config(self.h, int(path), int(tile)), self.h)

    def positions(self, m=0):
        """Vertex positions of member m, shape (n_vertex_slots, 2)."""
        nv = self.sizes(m)[0]
        out = np.empty(2 * nv)
        _check(lib().sevc_get_positions(self.h, int(m), out.ctypes.data_as(C.POINTER(C.c_double))), self.h)
        return out.reshape(-1, 2)

    def request_diagnostics(self):
        _check(lib().sevc_request_diagnostics(self.h), self.h)

    def checkpoint(self, slot=0):
        """Keep the whole device state in checkpoint `slot` (device memory)."""
        _check(lib().sevc_checkpoint(self.h, int(slot)), self.h)

    def restore(self, slot=0):
        """Put the state back as it was at checkpoint(slot)."""
        _check(lib().sevc_restore(self.h, int(slot)), self.h)

    def set_mesh(self, m, mesh):
        """Restore member m's whole mesh (a HostMesh of the member's sizes)."""
        _check(lib().sevc_set_member_mesh(self.h, int(m), C.byref(mesh.struct)), self.h)

    def set_params(self, m=0, pos=None, g_SE=None, gb_scaling=-1.0, mobility=-1.0):
        P = C.POINTER(C.c_double)
        pp = None if pos is None else np.ascontiguousarray(pos, dtype=np.float64)
        sp = None if g_SE is None else np.ascontiguousarray(g_SE, dtype=np.float64)
        _check(lib().sevc_set_member_params(self.h, int(m), None if pp is None else pp.ctypes.data_as(P),
                                            None if sp is None else sp.ctypes.data_as(P), float(gb_scaling),
                                            float(mobility)), self.h)

    def window_forward(self, steps, ckpt_every=0):
        """Topology-fixed window forward (analytic model, events off, fixed
        dt) on every member; checkpoints every ckpt_every steps (0: all)."""
        _check(lib().sevc_window_forward(self.h, int(steps), int(ckpt_every)), self.h)

    def adjoint(self, dL_dpos=None):
        """Reverse pass over the last window.  dL_dpos: one array per member
        (2 per vertex slot), a single array for a one-member state, or None
        for the device seed of area_histogram(seed=True).
        Returns, per member, (d_pos, d_SE, d_gb_scaling, d_mobility)."""
        if isinstance(dL_dpos, np.ndarray):
            dL_dpos = [dL_dpos]
        P = C.POINTER(C.c_double)
        outs = []
        arr_g = (_abi.Grads * self.n_members)()
        arr_in = None
        if dL_dpos is not None:
            ins = [np.ascontiguousarray(g, dtype=np.float64).reshape(-1) for g in dL_dpos]
            arr_in = (P * self.n_members)()
        for m in range(self.n_members):
            nv, ne, ng, S = self.sizes(m)
            if arr_in is not None:
                if ins[m].size != 2 * nv:
                    raise ValueError(f"member {m}: dL_dpos has {ins[m].size} values, expected {2 * nv}")
                arr_in[m] = ins[m].ctypes.data_as(P)
            dp, ds = np.zeros(2 * nv), np.zeros(ng)
            outs.append((dp, ds))
            arr_g[m].d_pos = dp.ctypes.data_as(P)
            arr_g[m].d_SE = ds.ctypes.data_as(P)
        _check(lib().sevc_adjoint_window(self.h, arr_in, arr_g), self.h)
        return [(dp, ds, arr_g[m].d_gb_scaling, arr_g[m].d_mobility) for m, (dp, ds) in enumerate(outs)]

    def enable_timing(self, on=True):
        """0/False off, 1/True every kernel, 2 only the velocity passes,
        translation and t_ext (HIP events on the state's stream)."""
        _check(lib().sevc_enable_kernel_timing(self.h, int(on)), self.h)

    def set_timing_sample(self, every):
        """Time the kernels of one step in `every` (sevc_step)."""
        _check(lib().sevc_set_timing_sample(self.h, int(every)), self.h)

    def kernel_times(self):
        n = 64
        names = (C.c_char * 48 * n)()
        ms = (C.c_double * n)()
        cnt = (C.c_int64 * n)()
        k = lib().sevc_kernel_times(self.h, n, C.cast(names, C.c_void_p), ms, cnt)
        return {bytes(names[i]).split(b"\0")[0].decode(): (ms[i], cnt[i]) for i in range(k)}

    def close(self):
        if getattr(self, "h", None):
            lib().sevc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def selftest_math(op, x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros_like(x)
    P = C.POINTER(C.c_double)
    _check(lib().sevc_selftest_math(int(op), x.size, x.ctypes.data_as(P), y.ctypes.data_as(P),
                                    out.ctypes.data_as(P)))
    return out
