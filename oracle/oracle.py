"""ctypes front end of the FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by the cpu_baseline leg of bench.py, never by the
trimem_amd package.  The oracle restates bio-phys/trimem's arithmetic
(src/mesh_util.h, src/mesh_properties.cpp, src/kernel.h, src/energy.cpp,
src/mesh_tether.h, src/mesh_repulsion.h, src/nlists/, src/flips.cpp,
src/flip_utils.cpp, src/util.cpp) on the CPU in FP64 with OpenMP.

Parity status of each function (how it is pinned) is listed in the test
modules tests/test_oracle_*.py; everything here is pinned by at least one
check that does not re-type the oracle's own formula, except the OpenMesh
circulation start and the OpenMesh flip rewiring, which no reference test
constrains (parity unpinned, see tests/test_oracle_flips.py).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class orc_params(ctypes.Structure):
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
        ("rep_exclusion", ctypes.c_int), ("pad_", ctypes.c_int),
        ("cont_delta", ctypes.c_double), ("cont_lambda", ctypes.c_double),
    ]


def build():
    """Compile liboracle.so from oracle.cpp (make)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P, I, D, L_ = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        sig = {
            "orc_last_error": (ctypes.c_char_p, []),
            "orc_num_threads": (I, []),
            "orc_mesh_new": (P, [dp, I, ip, I]),
            "orc_mesh_free": (None, [P]),
            "orc_mesh_nv": (I, [P]), "orc_mesh_ne": (I, [P]), "orc_mesh_nf": (I, [P]),
            "orc_mesh_get_points": (None, [P, dp]),
            "orc_mesh_set_points": (None, [P, dp]),
            "orc_mesh_fv_indices": (None, [P, ip]),
            "orc_mesh_topology": (None, [P, ip, ip, ip, ip, ip]),
            "orc_mesh_is_flip_ok": (I, [P, I]),
            "orc_mesh_flip": (None, [P, I]),
            "orc_mesh_next_halfedge": (I, [P, I]),
            "orc_edge_length": (D, [P, I]), "orc_edge_length_grad": (None, [P, I, dp]),
            "orc_face_area": (D, [P, I]), "orc_face_area_grad": (None, [P, I, dp]),
            "orc_face_volume": (D, [P, I]), "orc_face_volume_grad": (None, [P, I, dp]),
            "orc_dihedral_angle": (D, [P, I]), "orc_dihedral_angle_grad": (None, [P, I, dp]),
            "orc_energy_new": (P, [P, ctypes.POINTER(orc_params)]),
            "orc_energy_free": (None, [P]),
            "orc_properties": (I, [P, P, dp]),
            "orc_vertex_properties": (I, [P, P, dp]),
            "orc_energy": (D, [P, P]),
            "orc_energy_props": (D, [P, dp]),
            "orc_gradient": (I, [P, P, dp]),
            "orc_update_reference_properties": (None, [P]),
            "orc_update_repulsion": (I, [P, P]),
            "orc_initial_props": (None, [P, dp]),
            "orc_get_params": (None, [P, ctypes.POINTER(orc_params)]),
            "orc_fd_gradient": (I, [P, P, dp, D]),
            "orc_nlist_new": (P, [P, ctypes.POINTER(orc_params)]),
            "orc_nlist_free": (None, [P]),
            "orc_nlist_distance_matrix": (L_, [P, P, D, dp, ip, ip, L_]),
            "orc_nlist_point_distances": (L_, [P, P, I, D, dp, ip, L_]),
            "orc_energy_nlist": (P, [P]),
            "orc_flip_serial_injected": (I, [P, P, I, ip, dp, I, ip]),
            "orc_edge_vertex_properties": (I, [P, P, I, dp]),
            "orc_area": (D, [P]), "orc_edges_length": (D, [P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _err():
    return lib().orc_last_error().decode()


def num_threads():
    return lib().orc_num_threads()


PROP_NAMES = ("area", "volume", "curvature", "bending", "tethering", "repulsion")


class Params:
    """Flat mirror of trimem EnergyParams (src/params.h:19-71) with its defaults."""

    def __init__(self, **kw):
        self.kappa_b = 0.0
        self.kappa_a = 0.0
        self.kappa_v = 0.0
        self.kappa_c = 0.0
        self.kappa_t = 0.0
        self.kappa_r = 0.0
        self.area_frac = 1.0
        self.volume_frac = 1.0
        self.curvature_frac = 1.0
        self.bond_lc0 = 1.0
        self.bond_lc1 = 0.0
        self.bond_a0 = 1.0
        self.bond_r = 2
        self.bond_type = 0
        self.rep_lc1 = 0.0
        self.rep_r = 2
        self.rep_nlist = 0
        self.rep_rlist = 0.1
        self.rep_exclusion = 2
        self.cont_delta = 0.0
        self.cont_lambda = 1.0
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(k)
            setattr(self, k, v)

    def c(self):
        p = orc_params()
        for name, _ in orc_params._fields_:
            if name != "pad_":
                setattr(p, name, getattr(self, name))
        return p


class Mesh:
    def __init__(self, points, cells):
        points = np.ascontiguousarray(points, dtype=np.float64).reshape((-1, 3))
        cells = np.ascontiguousarray(cells, dtype=np.int32).reshape((-1, 3))
        self._h = lib().orc_mesh_new(_dp(points), len(points), _ip(cells), len(cells))
        if not self._h:
            raise RuntimeError(_err())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_mesh_free(self._h)
            self._h = None

    def n_vertices(self):
        return lib().orc_mesh_nv(self._h)

    def n_edges(self):
        return lib().orc_mesh_ne(self._h)

    def n_faces(self):
        return lib().orc_mesh_nf(self._h)

    @property
    def x(self):
        out = np.empty((self.n_vertices(), 3))
        lib().orc_mesh_get_points(self._h, _dp(out))
        return out

    @x.setter
    def x(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape((self.n_vertices(), 3))
        lib().orc_mesh_set_points(self._h, _dp(v))

    def fv_indices(self):
        out = np.empty((self.n_faces(), 3), dtype=np.int32)
        lib().orc_mesh_fv_indices(self._h, _ip(out))
        return out

    def topology(self):
        nh = 2 * self.n_edges()
        to = np.empty(nh, np.int32)
        nx = np.empty(nh, np.int32)
        fc = np.empty(nh, np.int32)
        vh = np.empty(self.n_vertices(), np.int32)
        fh = np.empty(self.n_faces(), np.int32)
        lib().orc_mesh_topology(self._h, _ip(to), _ip(nx), _ip(fc), _ip(vh), _ip(fh))
        return dict(he_to=to, he_next=nx, he_face=fc, v_he=vh, f_he=fh)

    def is_flip_ok(self, e):
        return bool(lib().orc_mesh_is_flip_ok(self._h, int(e)))

    def flip(self, e):
        lib().orc_mesh_flip(self._h, int(e))

    def next_halfedge(self, he):
        return lib().orc_mesh_next_halfedge(self._h, int(he))

    # geometry primitives on halfedge he (mesh_util.h)
    def edge_length(self, he):
        return lib().orc_edge_length(self._h, he)

    def edge_length_grad(self, he):
        out = np.empty((2, 3))
        lib().orc_edge_length_grad(self._h, he, _dp(out))
        return out

    def face_area(self, he):
        return lib().orc_face_area(self._h, he)

    def face_area_grad(self, he):
        out = np.empty((3, 3))
        lib().orc_face_area_grad(self._h, he, _dp(out))
        return out

    def face_volume(self, he):
        return lib().orc_face_volume(self._h, he)

    def face_volume_grad(self, he):
        out = np.empty((3, 3))
        lib().orc_face_volume_grad(self._h, he, _dp(out))
        return out

    def dihedral_angle(self, he):
        return lib().orc_dihedral_angle(self._h, he)

    def dihedral_angle_grad(self, he):
        out = np.empty((4, 3))
        lib().orc_dihedral_angle_grad(self._h, he, _dp(out))
        return out

    def area(self):
        return lib().orc_area(self._h)

    def edges_length(self):
        return lib().orc_edges_length(self._h)

    def avg_tri_props(self):
        return self.area() / self.n_faces(), self.edges_length() / self.n_edges()


class NeighbourList:
    def __init__(self, mesh, params=None, handle=None, owner=None):
        if handle is not None:
            self._h = handle
            self._own = False
            self._owner = owner
        else:
            c = params.c()
            self._h = lib().orc_nlist_new(mesh._h, ctypes.byref(c))
            self._own = True
            if not self._h:
                raise RuntimeError(_err())

    def __del__(self):
        if getattr(self, "_own", False) and self._h:
            lib().orc_nlist_free(self._h)
            self._h = None

    def distance_matrix(self, mesh, dmax):
        n = lib().orc_nlist_distance_matrix(self._h, mesh._h, dmax, None, None, None, 0)
        if n < 0:
            raise RuntimeError(_err())
        d = np.empty(n)
        i = np.empty(n, np.int32)
        j = np.empty(n, np.int32)
        lib().orc_nlist_distance_matrix(self._h, mesh._h, dmax, _dp(d), _ip(i), _ip(j), n)
        return d, i, j

    def point_distances(self, mesh, pid, dmax):
        n = lib().orc_nlist_point_distances(self._h, mesh._h, pid, dmax, None, None, 0)
        if n < 0:
            raise RuntimeError(_err())
        d = np.empty((n, 3))
        j = np.empty(n, np.int32)
        lib().orc_nlist_point_distances(self._h, mesh._h, pid, dmax, _dp(d), _ip(j), n)
        return d, j


class EnergyManager:
    def __init__(self, mesh, params):
        self._c = params.c()
        self._h = lib().orc_energy_new(mesh._h, ctypes.byref(self._c))
        if not self._h:
            raise RuntimeError(_err())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_energy_free(self._h)
            self._h = None

    def properties(self, mesh):
        out = np.empty(6)
        if lib().orc_properties(self._h, mesh._h, _dp(out)) != 0:
            raise RuntimeError(_err())
        return out

    def vertex_properties(self, mesh):
        out = np.empty((mesh.n_vertices(), 6))
        if lib().orc_vertex_properties(self._h, mesh._h, _dp(out)) != 0:
            raise RuntimeError(_err())
        return out

    def energy(self, mesh):
        return lib().orc_energy(self._h, mesh._h)

    def energy_props(self, props):
        props = np.ascontiguousarray(props, dtype=np.float64)
        return lib().orc_energy_props(self._h, _dp(props))

    def gradient(self, mesh):
        out = np.empty((mesh.n_vertices(), 3))
        if lib().orc_gradient(self._h, mesh._h, _dp(out)) != 0:
            raise RuntimeError(_err())
        return out

    def fd_gradient(self, mesh, eps):
        out = np.empty((mesh.n_vertices(), 3))
        if lib().orc_fd_gradient(self._h, mesh._h, _dp(out), eps) != 0:
            raise RuntimeError(_err())
        return out

    def update_reference_properties(self):
        lib().orc_update_reference_properties(self._h)

    def update_repulsion(self, mesh):
        if lib().orc_update_repulsion(self._h, mesh._h) != 0:
            raise RuntimeError(_err())

    def initial_props(self):
        out = np.empty(6)
        lib().orc_initial_props(self._h, _dp(out))
        return out

    def params(self):
        p = orc_params()
        lib().orc_get_params(self._h, ctypes.byref(p))
        return p

    def nlist(self):
        return NeighbourList(None, handle=lib().orc_energy_nlist(self._h), owner=self)

    def flip_serial_injected(self, mesh, proposals, uniforms):
        proposals = np.ascontiguousarray(proposals, dtype=np.int32)
        uniforms = np.ascontiguousarray(uniforms, dtype=np.float64)
        used = ctypes.c_int(0)
        acc = lib().orc_flip_serial_injected(self._h, mesh._h, len(proposals), _ip(proposals),
                                             _dp(uniforms), len(uniforms), ctypes.byref(used))
        if acc < 0:
            raise RuntimeError(_err())
        return acc, used.value

    def edge_vertex_properties(self, mesh, e):
        out = np.empty(6)
        if lib().orc_edge_vertex_properties(self._h, mesh._h, int(e), _dp(out)) != 0:
            raise RuntimeError(_err())
        return out
