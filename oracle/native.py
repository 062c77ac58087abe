"""ctypes binding of the oracle's C restatement (oracle/_build/libpetch_oracle.so).
TEST INFRASTRUCTURE ONLY — see oracle/petch_oracle.c."""
import ctypes
import os

import numpy as np

from . import build as _build

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.LIB):
            _build.build()
        _lib = ctypes.CDLL(_build.LIB)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().oracle_philox(_p(c), _p(k), _p(out))
    return out


def marching_cubes(phi, dx):
    """(verts float32 (nv,3), faces int32 (nf,3)) of the oracle MC."""
    phi = np.ascontiguousarray(phi, np.float32)
    nx, ny, nz = phi.shape
    L = lib()
    L.oracle_marching_cubes.restype = ctypes.c_int
    nv = ctypes.c_longlong()
    nf = ctypes.c_longlong()
    args = [_p(phi), ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz), ctypes.c_double(dx)]
    L.oracle_marching_cubes(*args, None, None, ctypes.c_longlong(0), ctypes.c_longlong(0),
                            ctypes.byref(nv), ctypes.byref(nf))
    verts = np.zeros((max(nv.value, 1), 3), np.float32)
    faces = np.zeros((max(nf.value, 1), 3), np.int32)
    rc = L.oracle_marching_cubes(*args, _p(verts), _p(faces), ctypes.c_longlong(nv.value),
                                 ctypes.c_longlong(nf.value), ctypes.byref(nv), ctypes.byref(nf))
    assert rc == 0
    return verts[:nv.value], faces[:nf.value]


def mc_cell(f8):
    f = np.ascontiguousarray(f8, np.float32)
    tri = np.zeros(30, np.int32)
    n = lib().oracle_mc_cell(_p(f), _p(tri))
    return tri[:3 * n].reshape(-1, 3)


def closest_hit_brute(verts, faces, orig, dirs, tmax=1.0e6):
    v = np.ascontiguousarray(verts, np.float32)
    f = np.ascontiguousarray(faces, np.int32)
    o = np.ascontiguousarray(orig, np.float32)
    d = np.ascontiguousarray(dirs, np.float32)
    n = len(o)
    face = np.zeros(n, np.int32)
    t = np.zeros(n, np.float64)
    lib().oracle_closest_hit_brute(_p(v), _p(f), ctypes.c_int(len(f)), _p(o), _p(d),
                                   ctypes.c_int(n), ctypes.c_double(tmax), _p(face), _p(t))
    return face, t


def closest_point_brute(verts, faces, pts):
    v = np.ascontiguousarray(verts, np.float32)
    f = np.ascontiguousarray(faces, np.int32)
    p = np.ascontiguousarray(pts, np.float32)
    n = len(p)
    face = np.zeros(n, np.int32)
    dist = np.zeros(n, np.float64)
    lib().oracle_closest_point_brute(_p(v), _p(f), ctypes.c_int(len(f)), _p(p), ctypes.c_int(n),
                                     _p(face), _p(dist))
    return face, dist
