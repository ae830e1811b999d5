"""GPU: tm_mesh_sum_noalloc, the allocation-free variant of tm_area and
tm_edges_length (util.cpp:9-35).

Called through the C ABI on two meshes in a row, the second smaller than
the first (the scratch array keeps the first call's partials and counter),
it returns bit for bit what tm_area / tm_edges_length return: both run the
same reduction kernel over the same blocks.  Its error paths (a mesh not
yet on the device, a bad selector) return TM_E_INVALID.
"""
import ctypes

import numpy as np
import pytest

from trimem_amd import capi, meshes

pytestmark = pytest.mark.gpu


def hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


def test_mesh_sum_noalloc_matches_area_and_edges_length():
    import trimem_amd.core as core
    assert core.device_count() > 0
    L = capi.lib()
    L.tm_mesh_sum_noalloc.restype = ctypes.c_int
    L.tm_mesh_sum_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.tm_area.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    L.tm_edges_length.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    H = hip()
    d_out = ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(d_out), 8) == 0
    try:
        for level, seed in ((6, 1), (3, 2)):            # the second call on a smaller mesh
            v, f, _ = meshes.jittered_icosphere(level, seed, rel_sigma=0.05)
            m = core.TriMesh(v, f)
            h = ctypes.c_void_p(m._handle())
            for what, ref_fn in ((0, L.tm_area), (1, L.tm_edges_length)):
                ref = ctypes.c_double()
                capi.check(ref_fn(h, ctypes.byref(ref)))    # puts the mesh on the device
                capi.check(L.tm_mesh_sum_noalloc(h, what, d_out))
                capi.check(L.tm_sync())
                got = ctypes.c_double()
                assert H.hipMemcpy(ctypes.byref(got), d_out, 8, 2) == 0    # device to host
                print(f"level {level} what {what}: {got.value!r} vs {ref.value!r}", flush=True)
                assert got.value == ref.value
            # the reference value from the mesh as numpy sees it
            tri = v[f]
            a = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1).sum()
            capi.check(L.tm_mesh_sum_noalloc(h, 0, d_out))
            capi.check(L.tm_sync())
            got = ctypes.c_double()
            assert H.hipMemcpy(ctypes.byref(got), d_out, 8, 2) == 0
            assert abs(got.value - a) <= 1e-12 * a
        # a mesh whose arrays are not on the device yet, and a bad selector
        v, f = meshes.icosphere(1)
        fresh = core.TriMesh(v, f)
        assert L.tm_mesh_sum_noalloc(ctypes.c_void_p(fresh._handle()), 0, d_out) == -1
        assert L.tm_mesh_sum_noalloc(h, 2, d_out) == -1
    finally:
        H.hipFree(d_out)
