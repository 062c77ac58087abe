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


# ------------------------------------------------------------------ traces
class Box(ctypes.Structure):
    _fields_ = [("Lx", ctypes.c_float), ("Ly", ctypes.c_float), ("Lz", ctypes.c_float),
                ("z_src", ctypes.c_float), ("periodic", ctypes.c_int)]


class Source(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("sigma", ctypes.c_float), ("exponent", ctypes.c_float)]


class Ion(ctypes.Structure):
    _fields_ = [(k, ctypes.c_float) for k in ("meanE", "sigmaE", "A_sp", "Eth_sp", "B_sp", "A_ie",
                                              "Eth_ie", "A_p", "Eth_p", "inflect", "minang",
                                              "n_l", "A_en", "minE", "thrmin", "thrmax")]


class OracleBvh:
    """The oracle's own CPU BVH over a fixed mesh (median split; closest hit in float64)."""

    def __init__(self, verts, faces):
        L = lib()
        L.oracle_bvh_build.restype = ctypes.c_void_p
        self.v = np.ascontiguousarray(verts, np.float32)
        self.f = np.ascontiguousarray(faces, np.int32)
        self.nf = len(self.f)
        self.h = ctypes.c_void_p(L.oracle_bvh_build(_p(self.v), _p(self.f), ctypes.c_int(self.nf)))

    def __del__(self):
        try:
            lib().oracle_bvh_free(self.h)
        except Exception:
            pass

    def closest_hit(self, orig, dirs, tmax=1.0e6):
        o = np.ascontiguousarray(orig, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = len(o)
        face = np.zeros(n, np.int32)
        t = np.zeros(n, np.float64)
        lib().oracle_bvh_closest_hit(self.h, _p(o), _p(d), ctypes.c_int(n), ctypes.c_double(tmax),
                                     _p(face), _p(t))
        return face, t

    def closest_point(self, pts):
        p = np.ascontiguousarray(pts, np.float32)
        n = len(p)
        face = np.zeros(n, np.int32)
        dist = np.zeros(n, np.float64)
        lib().oracle_bvh_closest_point(self.h, _p(p), ctypes.c_int(n), _p(face), _p(dist))
        return face, dist

    def trace_neutral_rr(self, box, src, bare, beta, n, seed):
        bare = np.ascontiguousarray(bare, np.float32)
        out = np.zeros(self.nf, np.float64)
        lib().oracle_trace_neutral_rr(self.h, ctypes.byref(box), ctypes.byref(src), _p(bare),
                                      ctypes.c_float(beta), ctypes.c_int(n),
                                      ctypes.c_uint64(seed), _p(out))
        return out

    def trace_neutral_cap(self, box, src, bare, beta, n_reemit, n, seed):
        """Per-face arrivals (float64 (F,)) of oracle_trace_neutral_cap."""
        bare = np.ascontiguousarray(bare, np.float32)
        out = np.zeros(max(self.nf, 1), np.float64)
        lib().oracle_trace_neutral_cap(self.h, ctypes.byref(box), ctypes.byref(src), _p(bare),
                                       ctypes.c_float(beta), ctypes.c_int(n_reemit),
                                       ctypes.c_int(n), ctypes.c_uint64(seed), _p(out))
        return out[:self.nf]

    def trace_redeposition(self, cen, gn, area, rate, s_redep, n, seed):
        """Redeposited flux density (float64 (F,)) of oracle_trace_redeposition."""
        c = np.ascontiguousarray(cen, np.float32)
        g = np.ascontiguousarray(gn, np.float32)
        a = np.ascontiguousarray(area, np.float32)
        r = np.ascontiguousarray(rate, np.float32)
        out = np.zeros(max(self.nf, 1), np.float64)
        lib().oracle_trace_redeposition(self.h, _p(c), _p(g), _p(a), _p(r), ctypes.c_float(s_redep),
                                        ctypes.c_int(n), ctypes.c_uint64(seed), _p(out))
        return out[:self.nf]

    def trace_neutral_rr_paths(self, box, src, bare, beta, n, seed):
        """trace_neutral_rr plus each ray's hit count (int32) and deposited weight (float64)."""
        bare = np.ascontiguousarray(bare, np.float32)
        out = np.zeros(self.nf, np.float64)
        nhit = np.zeros(n, np.int32)
        wsum = np.zeros(n, np.float64)
        lib().oracle_trace_neutral_rr_paths(self.h, ctypes.byref(box), ctypes.byref(src),
                                            _p(bare), ctypes.c_float(beta), ctypes.c_int(n),
                                            ctypes.c_uint64(seed), _p(out), _p(nhit), _p(wsum))
        return out, nhit, wsum

    def trace_neutral_rr_adjoint(self, box, src, bare, beta, n, seed, gfl):
        """d loss / d bare of trace_neutral_rr for gfl = d loss / d fl, paths and decisions fixed."""
        bare = np.ascontiguousarray(bare, np.float32)
        g = np.ascontiguousarray(gfl, np.float64)
        out = np.zeros(self.nf, np.float64)
        lib().oracle_trace_neutral_rr_adjoint(self.h, ctypes.byref(box), ctypes.byref(src),
                                              _p(bare), ctypes.c_float(beta), ctypes.c_int(n),
                                              ctypes.c_uint64(seed), _p(g), _p(out))
        return out

    def trace_neutral_rr_frozen(self, box, src, bare_dec, bare_w, beta, n, seed):
        """trace_neutral_rr with paths and decisions from bare_dec, weights (float64) from
        bare_w."""
        a = np.ascontiguousarray(bare_dec, np.float32)
        b = np.ascontiguousarray(bare_w, np.float32)
        out = np.zeros(self.nf, np.float64)
        lib().oracle_trace_neutral_rr_frozen(self.h, ctypes.byref(box), ctypes.byref(src), _p(a),
                                             _p(b), ctypes.c_float(beta), ctypes.c_int(n),
                                             ctypes.c_uint64(seed), _p(out))
        return out

    def trace_ion_yield(self, box, src, ion, n, seed):
        out = np.zeros((5, self.nf), np.float64)
        lib().oracle_trace_ion_yield(self.h, ctypes.byref(box), ctypes.byref(src),
                                     ctypes.byref(ion), ctypes.c_int(n), ctypes.c_uint64(seed),
                                     _p(out))
        return out

    def trace_legacy(self, box, src, sticking, n_reemit, specular, cos_thr, eta, n, seed):
        out = np.zeros((2, self.nf), np.float64)
        lib().oracle_trace_legacy(self.h, ctypes.byref(box), ctypes.byref(src),
                                  ctypes.c_float(sticking), ctypes.c_int(n_reemit),
                                  ctypes.c_int(specular), ctypes.c_float(cos_thr),
                                  ctypes.c_float(eta), ctypes.c_int(n), ctypes.c_uint64(seed),
                                  _p(out))
        return out

    def trace_ff(self, box, centroids, gas_normals, n_ff, seed):
        c = np.ascontiguousarray(centroids, np.float32)
        g = np.ascontiguousarray(gas_normals, np.float32)
        out = np.zeros(len(c) * n_ff, np.int32)
        lib().oracle_trace_ff(self.h, ctypes.byref(box), _p(c), _p(g), ctypes.c_int(len(c)),
                              ctypes.c_int(n_ff), ctypes.c_uint64(seed), _p(out))
        return out


def redep_faces(rate, area, n, seed):
    """The emitting face of each of n redeposition rays (int32 (n,)), oracle_redep_faces."""
    r = np.ascontiguousarray(rate, np.float32)
    a = np.ascontiguousarray(area, np.float32)
    out = np.zeros(n, np.int32)
    lib().oracle_redep_faces(_p(r), _p(a), ctypes.c_int(len(r)), ctypes.c_int(n),
                             ctypes.c_uint64(seed), _p(out))
    return out


def gen_source(box, src, n, seed):
    o = np.zeros((n, 3), np.float32)
    d = np.zeros((n, 3), np.float32)
    lib().oracle_gen_source(ctypes.byref(box), ctypes.byref(src), ctypes.c_int(n),
                            ctypes.c_uint64(seed), _p(o), _p(d))
    return o, d
