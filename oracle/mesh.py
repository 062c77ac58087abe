"""Mesh-side oracle (FP64 numpy) — TEST INFRASTRUCTURE ONLY, never imported by the product.

Restates, from the reference src/petch/threed.py:
  mesh_attrs      centroids / areas            threed.py:96-99 (and 65-68)
  face_normals    _face_normals_k / host path  threed.py:228-234, 773-775
  edge_adjacency  _edge_adjacency (numpy)      threed.py:118-125
  smooth_flux     smooth_flux                  threed.py:129-148
  smooth_flux_dev smooth_flux_dev              threed.py:195-214 (per-call alpha blend)
  norm_clip       _norm_clip / host clip       threed.py:165-170, 837, 889
  cos_ratio       _cos_k / host np.where       threed.py:179-185, 821
  gas_normals     _gas_normals                 threed.py:1553-1568
plus mesh invariants used to pin the marching-cubes rule (closedness, orientation, Euler
characteristic, enclosed volume).
"""
import numpy as np


def mesh_attrs(verts, faces):
    v = np.asarray(verts, np.float64)[np.asarray(faces)]
    cen = v.mean(axis=1)
    cross = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    area = 0.5 * np.linalg.norm(cross, axis=1)
    return cen, area, cross


def face_normals(verts, faces):
    _, _, cross = mesh_attrs(verts, faces)
    n = np.linalg.norm(cross, axis=1, keepdims=True)
    return np.where(n > 0, cross / np.where(n > 0, n, 1.0), 0.0)


def edge_adjacency(faces):
    """threed.py:118-125: sort the 3F (sorted) edge keys stably and pair consecutive equal keys."""
    faces = np.asarray(faces)
    F = len(faces)
    if F == 0:
        return np.zeros((0, 2), np.int64)
    e = np.sort(faces[:, [[0, 1], [1, 2], [0, 2]]].reshape(-1, 2), axis=1)
    nv = int(faces.max()) + 1
    key = e[:, 0].astype(np.int64) * nv + e[:, 1]
    fid = np.repeat(np.arange(F), 3)
    order = np.argsort(key, kind='stable')
    key_s = key[order]
    fid_s = fid[order]
    idx = np.where(key_s[:-1] == key_s[1:])[0]
    return np.stack([fid_s[idx], fid_s[idx + 1]], axis=1)


def pairs_from_nbr(nbr):
    """The unordered pair set {min(f,g), max(f,g)} of a per-face neighbour table."""
    nbr = np.asarray(nbr).reshape(-1, 3)
    f = np.repeat(np.arange(len(nbr)), 3)
    g = nbr.ravel()
    m = g >= 0
    p = np.stack([np.minimum(f[m], g[m]), np.maximum(f[m], g[m])], 1)
    return np.unique(p, axis=0)


def canonical_pairs(pairs):
    p = np.asarray(pairs).reshape(-1, 2)
    p = np.stack([p.min(1), p.max(1)], 1)
    return np.unique(p, axis=0)


def smooth_flux(flux, normals, pairs, n_iter=1, alpha=1.0):
    """threed.py:129-148."""
    pairs = np.asarray(pairs)
    if len(pairs) == 0 or n_iter <= 0 or alpha <= 0.0:
        return np.asarray(flux, np.float64)
    raw = np.asarray(flux, np.float64)
    out = raw.copy()
    i, j = pairs[:, 0], pairs[:, 1]
    w = np.maximum(0.0, np.sum(normals[i] * normals[j], axis=1))
    for _ in range(n_iter):
        num = out.copy()
        den = np.ones(len(out))
        np.add.at(num, i, out[j] * w)
        np.add.at(den, i, w)
        np.add.at(num, j, out[i] * w)
        np.add.at(den, j, w)
        out = num / den
    return out if alpha >= 1.0 else (1.0 - alpha) * raw + alpha * out


def smooth_flux_dev(src, normals, pairs, alpha=1.0):
    """threed.py:195-214: one scatter + divide, blended with its own input when alpha < 1."""
    src = np.asarray(src, np.float64)
    if alpha <= 0.0:
        return src.copy()
    out = smooth_flux(src, normals, pairs, 1, 1.0)
    return (1.0 - alpha) * src + alpha * out if alpha < 1.0 else out


def norm_clip(fl, area, scale, lo, hi):
    """threed.py:165-170: clamp((fl/A)*scale, lo, hi)."""
    return np.clip((np.asarray(fl, np.float64) / np.asarray(area, np.float64)) * scale, lo, hi)


def cos_ratio(fl, ang):
    """threed.py:179-185 / 821: ang/fl where fl > 0 else 0."""
    fl = np.asarray(fl, np.float64)
    return np.where(fl > 0, np.asarray(ang, np.float64) / np.where(fl > 0, fl, 1.0), 0.0)


def gas_normals(verts, faces, centroids, phi, dx):
    """threed.py:1553-1568."""
    vv = np.asarray(verts, np.float64)[np.asarray(faces)]
    fn = np.cross(vv[:, 1] - vv[:, 0], vv[:, 2] - vv[:, 0])
    fn = fn / (np.linalg.norm(fn, axis=1, keepdims=True) + 1e-12)
    gx, gy, gz = np.gradient(np.asarray(phi, np.float64), dx)
    c = np.asarray(centroids, np.float64)
    ix = np.clip((c[:, 0] / dx).round().astype(int), 0, gx.shape[0] - 1)
    iy = np.clip((c[:, 1] / dx).round().astype(int), 0, gx.shape[1] - 1)
    iz = np.clip((c[:, 2] / dx).round().astype(int), 0, gx.shape[2] - 1)
    into_gas = -np.stack([gx[ix, iy, iz], gy[ix, iy, iz], gz[ix, iy, iz]], axis=1)
    flip = np.sum(fn * into_gas, axis=1) < 0.0
    fn[flip] *= -1.0
    return fn


# ------------------------------------------------------------------ mesh invariants
def directed_edge_counts(faces):
    """Counts of each directed edge (a->b) over the faces' cyclic edges."""
    f = np.asarray(faces, np.int64)
    d = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    keys, cnt = np.unique(d, axis=0, return_counts=True)
    return keys, cnt


def euler_characteristic(faces):
    f = np.asarray(faces, np.int64)
    V = len(np.unique(f))
    E = len(canonical_pairs(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])))
    return V - E + len(f)


def enclosed_volume(verts, faces):
    """Divergence theorem: sum of signed tetrahedra (origin, v0, v1, v2) / 6."""
    v = np.asarray(verts, np.float64)[np.asarray(faces)]
    return np.einsum('ij,ij->i', v[:, 0], np.cross(v[:, 1], v[:, 2])).sum() / 6.0
