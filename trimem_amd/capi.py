"""ctypes declarations of the C ABI in include/trimem_hip.h.

Used by bench.py for device-resident timing and by the ABI tests; the
pybind11 module trimem_amd.core is the drop-in for trimem.core.  Loading fails
loudly if libtrimem_hip.so has not been built.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtrimem_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "trimem_hip.h")

TM_OK = 0
TM_PREC_FP64, TM_PREC_MIXED = 0, 1
TM_EVAL_PROPS, TM_EVAL_GRAD = 1, 2


class tm_params(ctypes.Structure):
    _fields_ = [
        ("kappa_b", ctypes.c_double), ("kappa_a", ctypes.c_double),
        ("kappa_v", ctypes.c_double), ("kappa_c", ctypes.c_double),
        ("kappa_t", ctypes.c_double), ("kappa_r", ctypes.c_double),
        ("area_frac", ctypes.c_double), ("volume_frac", ctypes.c_double),
        ("curvature_frac", ctypes.c_double),
        ("bond_lc0", ctypes.c_double), ("bond_lc1", ctypes.c_double),
        ("bond_a0", ctypes.c_double), ("bond_r", ctypes.c_int),
        ("bond_type", ctypes.c_int),
        ("rep_lc1", ctypes.c_double), ("rep_r", ctypes.c_int),
        ("rep_nlist", ctypes.c_int), ("rep_rlist", ctypes.c_double),
        ("rep_exclusion", ctypes.c_int), ("reserved", ctypes.c_int),
        ("cont_delta", ctypes.c_double), ("cont_lambda", ctypes.c_double),
    ]


class tm_tuning(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int) for k in
                ("rep_threads", "rep_batch", "pass2_threads", "rep_fused", "pair_cache", "pflip_window", "serial_window")]


class tm_props(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in
                ("area", "volume", "curvature", "bending", "tethering", "repulsion")]


def declared_functions():
    """Names of the functions declared in include/trimem_hip.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tm_[a-z0-9_]+)\s*\(", text)))


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built (run python -m trimem_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    sig = {
        "tm_last_error": (ctypes.c_char_p, []),
        "tm_device_count": (I, [ip]),
        "tm_mesh_create": (I, [dp, I, ip, I, ctypes.POINTER(P)]),
        "tm_mesh_destroy": (I, [P]),
        "tm_mesh_counts": (I, [P, ip, ip, ip]),
        "tm_mesh_points": (dp, [P]),
        "tm_mesh_device_points": (dp, [P]),
        "tm_mesh_upload_points": (I, [P]),
        "tm_mesh_download_points": (I, [P]),
        "tm_energy_create": (I, [P, ctypes.POINTER(tm_params), ctypes.POINTER(P)]),
        "tm_energy_destroy": (I, [P]),
        "tm_energy_set_precision": (I, [P, I]),
        "tm_energy_energy_gradient": (I, [P, P, dp, dp]),
        "tm_energy_eval_device": (I, [P, P, I]),
        "tm_energy_device_gradient": (dp, [P]),
        "tm_energy_device_globals": (dp, [P]),
        "tm_energy_update_repulsion": (I, [P, P]),
        "tm_energy_get_tuning": (I, [P, ctypes.POINTER(tm_tuning)]),
        "tm_energy_set_tuning": (I, [P, ctypes.POINTER(tm_tuning)]),
        "tm_sync": (I, []),
        "tm_stream": (I, [ctypes.POINTER(P)]),
        "tm_mark_pass1": (I, [P, P]),
        "tm_energy_pflip_state": (I, [P, ctypes.POINTER(ctypes.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != TM_OK:
        raise RuntimeError(lib().tm_last_error().decode())
