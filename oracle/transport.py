"""Oracle (test infrastructure only) for the two deterministic neutral transports behind
Flags.neutral_transport: the 1-D Knudsen conductance tail (src/petch/knudsen.py) and the
discrete-ordinates grid march (src/petch/dda.py). Float64 numpy, except the DDA march, which
restates the reference's float32 Warp kernel (dda.py:26-95) one float32 operation at a time so
that the wall decisions of the HIP kernel can be compared exactly.

Pinned by tests/test_oracle_transport.py: the tridiagonal solve against scipy's banded solver,
the conductance profile of a uniform duct against its closed form, the slice geometry of boxes,
the nearest-slice rule, the march's direct flux on a flat field (exactly 1) and at the centre of a
long slot's floor (the view factor sin(theta0) of the opening), full sticking (no re-emission),
and the Aitken step on an exactly geometric sequence.
"""
import numpy as np

# ---------------------------------------------------------------- Knudsen (knudsen.py) --------


def slice_geometry(phi, zs, dx, sub_top, shape):
    """Per-slice gas area a(z) and sidewall perimeter p(z) (knudsen.py:24-71): a hole counts
    its gas cells (area) and the gas/solid transitions along x and y (perimeter, at least dx);
    a trench (running along y) takes the mean open x-width of the rows that have one, with a
    perimeter of 2 (the two long walls). Slices at or above sub_top stay 0."""
    nz = phi.shape[2]
    area = np.zeros(nz)
    perim = np.zeros(nz)
    gas = phi < 0.0
    for k in np.flatnonzero(np.asarray(zs) < sub_top):
        g = gas[:, :, k]
        if shape == "trench":
            w = g.sum(axis=0) * dx
            w = w[w > 0.0]
            if w.size:
                area[k] = float(w.mean())
                perim[k] = 2.0
        elif g.any():
            area[k] = float(g.sum()) * dx * dx
            tx = np.count_nonzero(g[1:, :] != g[:-1, :])
            ty = np.count_nonzero(g[:, 1:] != g[:, :-1])
            perim[k] = max(float(tx + ty) * dx, dx)
    return area, perim


def solve_tridiagonal(lower, diag, upper, rhs):
    """Thomas elimination with the reference's 1e-30 pivot guard (knudsen.py:74-90)."""
    n = len(diag)
    b = np.array(diag, float)
    d = np.array(rhs, float)
    for i in range(1, n):
        w = lower[i - 1] / max(b[i - 1], 1e-30)
        b[i] -= w * upper[i - 1]
        d[i] -= w * d[i - 1]
    x = np.zeros(n)
    if n:
        x[n - 1] = d[n - 1] / max(b[n - 1], 1e-30)
    for i in range(n - 2, -1, -1):
        x[i] = (d[i] - upper[i] * x[i + 1]) / max(b[i], 1e-30)
    return x


def conductance_profile(area, perim, zs, dx, sub_top, slice_loss, loss_scale):
    """Neutral concentration per slice (knudsen.py:93-151): the valid slices (below sub_top,
    with gas area and perimeter) from the opening down form a chain with free-molecular
    conductances C = (2/3) r a / dx, r = max(2a/p, dx); adjacent slices link by the harmonic
    mean 2 C0 C1 / (C0 + C1), non-adjacent ones not at all; the top slice couples to unit
    concentration through its own C; each slice loses loss_scale * slice_loss. Slices outside
    the chain read 1; the solution is clipped to [0, 1]. loss_scale: scalar or per slice."""
    nz = len(area)
    prof = np.ones(nz)
    ok = (np.asarray(zs) < sub_top) & (area > 0.0) & (perim > 0.0)
    if not ok.any():
        return prof
    idx = np.flatnonzero(ok)[::-1]
    a, p = area[idx], perim[idx]
    r = np.maximum(2.0 * a / np.maximum(p, 1e-30), dx)
    C = (2.0 / 3.0) * r * a / dx
    ls = np.asarray(loss_scale, float)
    sc = np.maximum(float(ls), 0.0) if ls.ndim == 0 else np.maximum(ls[idx], 0.0)
    diag = sc * np.maximum(slice_loss[idx], 0.0)
    m = len(idx)
    off = np.zeros(max(m - 1, 0))
    for i in range(m - 1):
        if abs(int(idx[i + 1]) - int(idx[i])) == 1:
            off[i] = 2.0 * C[i] * C[i + 1] / max(C[i] + C[i + 1], 1e-30)
    rhs = np.zeros(m)
    diag[0] += C[0]
    rhs[0] = C[0]
    for i in range(m - 1):
        diag[i] += off[i]
        diag[i + 1] += off[i]
    x = solve_tridiagonal(-off, np.maximum(diag, 1e-30), -off, rhs)
    prof[idx] = np.clip(x, 0.0, 1.0)
    return prof


def nearest_valid_slice(valid, iz):
    """Nearest valid slice of each iz, ties to the lower one; -1 without any (knudsen.py:154-163)."""
    valid = np.asarray(valid)
    iz = np.asarray(iz)
    if valid.size == 0:
        return np.full(iz.shape, -1, np.int64)
    pos = np.searchsorted(valid, iz)
    hi = valid[np.clip(pos, 0, valid.size - 1)]
    lo = valid[np.clip(pos - 1, 0, valid.size - 1)]
    return np.where(np.abs(hi - iz) < np.abs(lo - iz), hi, lo)


def front_loss_scale(zs, sub_top, W, wls, band_W, pass_frac, ar_pass):
    """Per-slice loss scale of the passivation-linked front loss (knudsen.py:166-182)."""
    ar = np.maximum((sub_top - np.asarray(zs)) / max(W, 1e-9), 0.0)
    g = np.exp(-np.maximum(ar - band_W, 0.0) / max(ar_pass, 1e-9))
    return max(float(wls), 0.0) * (pass_frac + (1.0 - pass_frac) * g)


def knudsen_face_flux(phi, zs, dx, sub_top, shape, centroids, gas_nz, areas, Ly, s_face,
                      wls=1.85, center=None, half_width=None, front_loss=False, band_W=8.0,
                      pass_frac=0.5, ar_pass=15.0):
    """Per-face neutral multiplier of the Knudsen tail (knudsen.py:185-241): faces below
    sub_top snap to their nearest valid slice; the floor faces (into-gas normal z > 0.7) inside
    the footprint (|x - cx| <= hw + 2dx for a trench, radius for a hole) feed 0.25 s A to their
    slice's sink (per unit length Ly for a trench); a face below sub_top reads its slice's
    concentration, any other face 1; clipped to [0, 8]."""
    nz = phi.shape[2]
    zs = np.asarray(zs, float)
    area, perim = slice_geometry(phi, zs, dx, sub_top, shape)
    valid = np.flatnonzero((zs < sub_top) & (area > 0.0) & (perim > 0.0))
    c = np.asarray(centroids, float)
    iz = np.clip(np.round((c[:, 2] - zs[0]) / dx).astype(np.int64), 0, nz - 1)
    snap = nearest_valid_slice(valid, iz)
    below = c[:, 2] < sub_top
    inside = np.ones(len(c), bool)
    if center is not None and half_width is not None:
        band = half_width + 2.0 * dx
        if shape == "trench":
            inside = np.abs(c[:, 0] - center[0]) <= band
        else:
            inside = np.hypot(c[:, 0] - center[0], c[:, 1] - center[1]) <= band
    loss = np.zeros(nz)
    sel = below & inside & (snap >= 0) & (np.asarray(gas_nz, float) > 0.7)
    if sel.any():
        np.add.at(loss, snap[sel],
                  0.25 * np.clip(np.asarray(s_face, float)[sel], 0.0, 1.0) * np.asarray(areas, float)[sel])
        if shape == "trench":
            loss /= max(float(Ly), 1e-30)
    scale = wls
    if front_loss:
        W = 2.0 * float(half_width) if half_width is not None else dx
        scale = front_loss_scale(zs, sub_top, W, wls, band_W, pass_frac, ar_pass)
    prof = conductance_profile(area, perim, zs, dx, sub_top, loss, scale)
    m = np.ones(len(c))
    use = (snap >= 0) & below
    m[use] = prof[snap[use]]
    return np.clip(m, 0.0, 8.0)


# ---------------------------------------------------------------- DDA (dda.py) ----------------


def fib_sphere(n):
    """n directions on the whole sphere, z uniform in (-1, 1) at (i + 1/2)/n, longitude advancing
    by the golden angle (dda.py:124-135); all weights 1."""
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(np.maximum(1.0 - z * z, 0.0))
    t = np.pi * (1.0 + 5.0 ** 0.5) * i
    return np.stack([r * np.cos(t), r * np.sin(t), z], 1), np.ones(n)


def _f(x):
    return np.float32(x)


def dda_gather(phi, dx, origins, normals, dirs, remit, z_top, src_top, max_steps, periodic_y):
    """The march of dda.py:26-95 in float32 (one operation, one rounding), vectorised over the
    origins for each direction: from o + 1.01 dx d, cell steps of dx d; at z >= z_top the ray
    carries src_top; where the trilinear phi (coordinates clamped to [0, n - 1.001]) is > 0 it
    carries remit of the first solid rounded cell among up to six half steps on (else of the
    rounded cell where it stopped); y wraps over ny dx with periodic_y. Returns
    sum_k acc_k c_k / max(sum_k acc_k, 1e-12) per origin, acc_k = max(n . d_k, 0) (the sums in
    float64, the reference's float32 sums differ in the last bits)."""
    phi = np.ascontiguousarray(phi, np.float32)
    nx, ny, nz = phi.shape
    flat = phi.ravel()
    o = np.asarray(origins, np.float32)
    nrm = np.asarray(normals, np.float32)
    D = np.asarray(dirs, np.float32)
    rem = np.asarray(remit, np.float32)
    dx32, zt, st = _f(dx), _f(z_top), _f(src_top)
    Ly = _f(_f(ny) * dx32)
    one, c101, half = _f(1.0), _f(1.01), _f(0.5)
    ux_max, uy_max, uz_max = _f(_f(nx) - _f(1.001)), _f(_f(ny) - _f(1.001)), _f(_f(nz) - _f(1.001))
    N = len(o)
    num = np.zeros(N)
    den = np.zeros(N)

    def node(i, j, k):
        return flat[(i * ny + j) * nz + k]

    def rounded(q, n):
        return np.clip((q / dx32 + half).astype(np.int64), 0, n - 1)

    for k in range(len(D)):
        d = D[k]
        acc = np.maximum(nrm[:, 0] * d[0] + nrm[:, 1] * d[1] + nrm[:, 2] * d[2], _f(0.0))
        acc = acc.astype(np.float32)
        den += acc
        act = np.flatnonzero(acc > _f(1e-6))
        if act.size == 0:
            continue
        sx, sy, sz = _f(d[0] * dx32), _f(d[1] * dx32), _f(d[2] * dx32)
        px = o[act, 0] + d[0] * dx32 * c101
        py = o[act, 1] + d[1] * dx32 * c101
        pz = o[act, 2] + d[2] * dx32 * c101
        px, py, pz = px.astype(np.float32), py.astype(np.float32), pz.astype(np.float32)
        contrib = np.zeros(act.size, np.float32)
        live = np.ones(act.size, bool)
        for _ in range(int(max_steps)):
            if not live.any():
                break
            L = np.flatnonzero(live)
            top = pz[L] >= zt
            contrib[L[top]] = st
            live[L[top]] = False
            W = L[~top]
            if W.size:
                u = np.clip(px[W] / dx32, _f(0.0), ux_max).astype(np.float32)
                v = np.clip(py[W] / dx32, _f(0.0), uy_max).astype(np.float32)
                w = np.clip(pz[W] / dx32, _f(0.0), uz_max).astype(np.float32)
                i0, j0, k0 = u.astype(np.int64), v.astype(np.int64), w.astype(np.int64)
                fu = (u - i0.astype(np.float32)).astype(np.float32)
                fv = (v - j0.astype(np.float32)).astype(np.float32)
                fw = (w - k0.astype(np.float32)).astype(np.float32)
                gu, gv, gw = one - fu, one - fv, one - fw
                p00 = node(i0, j0, k0) * gu + node(i0 + 1, j0, k0) * fu
                p10 = node(i0, j0 + 1, k0) * gu + node(i0 + 1, j0 + 1, k0) * fu
                p01 = node(i0, j0, k0 + 1) * gu + node(i0 + 1, j0, k0 + 1) * fu
                p11 = node(i0, j0 + 1, k0 + 1) * gu + node(i0 + 1, j0 + 1, k0 + 1) * fu
                pv = (p00 * gv + p10 * fv) * gw + (p01 * gv + p11 * fv) * fw
                wall = W[pv > _f(0.0)]
                if wall.size:
                    qx, qy, qz = px[wall].copy(), py[wall].copy(), pz[wall].copy()
                    got = np.zeros(wall.size, bool)
                    val = np.zeros(wall.size, np.float32)
                    for _e in range(6):
                        ix, iy, iz = rounded(qx, nx), rounded(qy, ny), rounded(qz, nz)
                        s = (~got) & (node(ix, iy, iz) > _f(0.0))
                        val[s] = rem[((ix * ny + iy) * nz + iz)[s]]
                        got |= s
                        qx = np.where(got, qx, qx + sx * half).astype(np.float32)
                        qy = np.where(got, qy, qy + sy * half).astype(np.float32)
                        qz = np.where(got, qz, qz + sz * half).astype(np.float32)
                    if (~got).any():
                        ix = rounded(px[wall], nx)
                        iy = rounded(py[wall], ny)
                        iz = rounded(pz[wall], nz)
                        val[~got] = rem[((ix * ny + iy) * nz + iz)[~got]]
                    contrib[wall] = val
                    live[wall] = False
            # every ray of the direction moves (finished ones harmlessly)
            px = (px + sx).astype(np.float32)
            py = (py + sy).astype(np.float32)
            pz = (pz + sz).astype(np.float32)
            if periodic_y:
                py = np.where(py < _f(0.0), py + Ly, py).astype(np.float32)
                py = np.where(py >= Ly, py - Ly, py).astype(np.float32)
        num[act] += acc[act].astype(np.float64) * contrib
    return num / np.maximum(den, 1e-12)


def dda_setup(phi, dx, zs, centroids, normals, areas, periodic_y):
    """Per-call constants of dda_neutral_flux (dda.py:155-176): the solid cell under each face
    (0.7 dx against its into-gas normal, rounded, clamped), the faces' area per cell, the cells
    seeded by a face, max_steps and z_top."""
    nx, ny, nz = phi.shape
    c = np.asarray(centroids, float)
    sp = c - 0.7 * dx * np.asarray(normals, float)
    inv = 1.0 / dx
    ix = np.clip((sp[:, 0] * inv + 0.5).astype(np.int64), 0, nx - 1)
    iy = np.clip((sp[:, 1] * inv + 0.5).astype(np.int64), 0, ny - 1)
    iz = np.clip((sp[:, 2] * inv + 0.5).astype(np.int64), 0, nz - 1)
    cell = (ix * ny + iy) * nz + iz
    area = np.zeros(nx * ny * nz)
    np.add.at(area, cell, np.asarray(areas, float))
    seeded = np.zeros(nx * ny * nz, bool)
    seeded[cell] = True
    z_top = float(zs[-1]) - 0.5 * dx
    max_steps = (int((z_top - float(zs[0])) / dx) + 4) * (3 if periodic_y else 1)
    return cell, area, seeded.reshape(nx, ny, nz), z_top, max_steps


def dda_remit(phi, cell, cell_area, seeded, areas, alpha, vf):
    """The per-cell re-emission field of one pass (dda.py:183-207): area-weighted mean of the
    faces' (1 - s) max(vf, 0) in their cells, then every solid unseeded cell with a seeded face
    neighbour (the six neighbours wrap around the grid) takes the mean of those neighbours."""
    shape = phi.shape
    R = np.zeros(int(np.prod(shape)))
    np.add.at(R, cell, np.asarray(areas, float) * alpha * np.maximum(vf, 0.0))
    R = (R / np.maximum(cell_area, 1e-12)).reshape(shape)
    need = (phi > 0.0) & ~seeded
    num = np.zeros(shape)
    cnt = np.zeros(shape)
    for ax in (0, 1, 2):
        for sh in (1, -1):
            s = np.roll(seeded, sh, axis=ax)
            num += np.where(s, np.roll(R, sh, axis=ax), 0.0)
            cnt += s
    fill = need & (cnt > 0)
    R[fill] = num[fill] / cnt[fill]
    return R.ravel()


def aitken(vf, prev1, prev2):
    """The geometric-tail extrapolation after the last pass (dda.py:212-218): per face
    r = d1/d0 (0 where |d0| <= 1e-12) clipped to [0, 0.995], vf + d1 r / (1 - r)."""
    d1 = vf - prev1
    d0 = prev1 - prev2
    big = np.abs(d0) > 1e-12
    r = np.clip(np.where(big, d1 / np.where(big, d0, 1.0), 0.0), 0.0, 0.995)
    return vf + d1 * r / (1.0 - r)


def dda_neutral_flux(phi, dx, zs, centroids, normals, s_face, areas, n_dir=64, n_reemit=12,
                     periodic_y=0):
    """Per-face neutral multiplier of the DDA transport (dda.py:138-220): the direct march from
    the sky, then n_reemit passes vf = direct + march(re-emission field), Aitken on the last
    three iterates, clipped to [0, 8]. The march runs from the centroids along 2 n_dir sphere
    directions (float32, as the reference hands them to its kernel)."""
    dirs, _ = fib_sphere(2 * n_dir)
    dirs = dirs.astype(np.float32)
    cell, carea, seeded, z_top, max_steps = dda_setup(phi, dx, zs, centroids, normals, areas,
                                                      periodic_y)
    c32 = np.asarray(centroids, np.float32)
    n32 = np.asarray(normals, np.float32)
    N = phi.size
    direct = dda_gather(phi, dx, c32, n32, dirs, np.zeros(N, np.float32), z_top, 1.0, max_steps,
                        periodic_y).astype(np.float32).astype(np.float64)
    vf = direct.copy()
    alpha = np.clip(1.0 - np.asarray(s_face, float), 0.0, 1.0)
    prev1 = prev2 = None
    for _ in range(max(1, int(n_reemit))):
        R = dda_remit(phi, cell, carea, seeded, areas, alpha, vf)
        prev2, prev1 = prev1, vf
        g = dda_gather(phi, dx, c32, n32, dirs, R.astype(np.float32), z_top, 0.0, max_steps,
                       periodic_y).astype(np.float32).astype(np.float64)
        vf = direct + g
    if prev2 is not None:
        vf = aitken(vf, prev1, prev2)
    return np.clip(np.maximum(vf, 0.0), 0.0, 8.0)
