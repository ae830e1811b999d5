"""GPU: tm_mark_pass1, the timing aid for the pass-1 kernel.

The next properties pass stamps the two events at the start and the end of
its pass-1 kernel, then forgets them; the pass does the same work either way,
so energy and gradient are bit-identical with and without the events, and
the stamped interval is a positive duration shorter than the whole call.
"""
import ctypes

import numpy as np
import pytest

from trimem_amd import capi
from trimem_amd import meshes as M

from test_gpu_pair_cache import params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    import trimem_amd.core as m
    assert m.device_count() > 0, "no GPU visible"
    return m


@pytest.fixture(scope="module")
def hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipEventDestroy.argtypes = [ctypes.c_void_p]
    h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    return h


def event(hip):
    e = ctypes.c_void_p()
    assert hip.hipEventCreate(ctypes.byref(e)) == 0
    return e


def elapsed_ms(hip, a, b):
    assert hip.hipEventSynchronize(b) == 0
    t = ctypes.c_float()
    assert hip.hipEventElapsedTime(ctypes.byref(t), a, b) == 0
    return float(t.value)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_pass1_marks_change_nothing(core, hip, precision):
    v, f, lbar = M.jittered_icosphere(4, seed=5, rel_sigma=0.02)
    m = core.TriMesh(v, f)
    e = core.EnergyManager(m, params(core, lbar, 2))
    e.precision = precision
    e0, g0 = e.energy(m), np.array(e.gradient(m))
    L = capi.lib()
    a, b = event(hip), event(hip)
    try:
        for _ in range(3):    # recording, then replays: both launch paths take the marks
            assert L.tm_mark_pass1(a, b) == 0
            g1 = np.array(e.gradient(m))
            t = elapsed_ms(hip, a, b)
            assert 0.0 < t < 100.0
            assert np.array_equal(g0, g1)
            assert e.energy(m) == e0
        # null events: nothing is stamped
        assert L.tm_mark_pass1(None, None) == 0
        assert np.array_equal(np.array(e.gradient(m)), g0)
    finally:
        hip.hipEventDestroy(a)
        hip.hipEventDestroy(b)
