"""The entry point that allocates nothing: sevc_check_meshes_device_noalloc
counts check_data's violations (utils.h:85-336) of meshes already in device
memory, with its scratch in a static device array.  Its counts must equal
those of sevc_check_mesh_device (the entry point that copies a host mesh to
the device) on two calls in a row, the second with smaller meshes, since
every call shares the one scratch array of the GPU."""
import ctypes as C

import numpy as np
import pytest
import torch

import sevc
from sevc._abi import Mesh
from _util import planted_corruptions, random_corruptions, ref2k_options

pytestmark = pytest.mark.gpu

_TOPO = ("v_bnd", "v_grn", "v_enabled", "b_vtx", "b_enabled", "g_vtx", "g_vlen", "g_enabled")
_CT = {np.int32: C.c_int32, np.uint8: C.c_uint8}


def _device_mesh(h, keep):
    """A Mesh whose topology arrays point at device copies (torch tensors
    kept alive in `keep`) of the host mesh h; the other arrays are NULL."""
    m = Mesh()
    m.n_vertices, m.n_boundaries, m.n_grains = h.nv, h.ne, h.ng
    m.max_vrt_per_grn, m.domain_bound = h.S, h.D
    for name in _TOPO:
        a = h[name]
        t = torch.from_numpy(a.copy()).to("cuda")
        keep.append(t)
        setattr(m, name, C.cast(C.c_void_p(t.data_ptr()), C.POINTER(_CT[a.dtype.type])))
    return m


def _noalloc_counts(hosts):
    keep = []
    arr = (Mesh * len(hosts))(*[_device_mesh(h, keep) for h in hosts])
    out = torch.full((len(hosts),), -7, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = sevc.lib().sevc_check_meshes_device_noalloc(arr, len(hosts), C.c_void_p(out.data_ptr()),
                                                      C.c_void_p(stream))
    assert rc == 0, sevc.lib().sevc_last_error(None)
    torch.cuda.synchronize()
    return out.cpu().numpy().tolist()


def _generated(n, gen_seed):
    pos, rid = sevc.voronoi(n, gen_seed)
    rng = np.random.default_rng(gen_seed)
    o = sevc.default_options()
    o.n_grains = n
    return sevc.build_mesh(o, pos, rid, rng.random(n), rng.random(n), n)


def test_noalloc_check_equals_the_allocating_entry_point():
    big = sevc.load_files(ref2k_options())
    first = [big] + [c for _, c, _ in planted_corruptions(big)] + random_corruptions(big, 9, seed=3)
    print(f"  call 1: {len(first)} meshes of 2,000 grains", flush=True)
    got = _noalloc_counts(first)
    want = [sevc.check_mesh_device(h) for h in first]
    assert got == want, (got, want)
    assert got[0] == 0 and all(g > 0 for g in got[1:7])

    small = _generated(60, 5)
    second = [small] + random_corruptions(small, 4, seed=11)
    print(f"  call 2: {len(second)} meshes of 60 grains", flush=True)
    got = _noalloc_counts(second)
    want = [sevc.check_mesh_device(h) for h in second]
    assert got == want, (got, want)
    assert got[0] == 0


def test_noalloc_rejects_more_meshes_than_its_scratch_holds():
    small = _generated(20, 3)
    keep = []
    dm = _device_mesh(small, keep)
    arr = (Mesh * 65)(*([dm] * 65))
    out = torch.zeros(65, dtype=torch.int32, device="cuda")
    rc = sevc.lib().sevc_check_meshes_device_noalloc(arr, 65, C.c_void_p(out.data_ptr()), None)
    assert rc == sevc.E_INVALID_ARG
    assert b"64" in sevc.lib().sevc_last_error(None)
    assert _noalloc_counts([small]) == [0]
