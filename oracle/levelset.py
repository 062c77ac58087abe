"""Level-set oracle (FP64 numpy) — TEST INFRASTRUCTURE ONLY, never imported by the product.

Restates, from the reference src/petch/threed.py:
  init_sdf      make_trench_3d            threed.py:32-59   (skfmm replaced by the exact distance;
                                                              skfmm's own output: parity unpinned)
  advect_3d     advect_3d                 threed.py:1248-1259
  advect_pinned advect_3d_gpu / CPU loop  threed.py:1295-1307, 1761-1763
  eik_solve     _eik_solve                threed.py:1421-1452
  eik_init      _eik_init                 threed.py:1455-1479
  eik_jacobi    _eik_jacobi               threed.py:1482-1512
  reinit_fsm    reinit_fsm                threed.py:1515-1537
  reinit_fsm_adjoint (no reference counterpart: its adjoint, pinned by central differences)
  depth3d / max_depth_3d                  threed.py:1798-1869
"""
import numpy as np


# ----------------------------------------------------------------------------- initial SDF
def grid_axes(Lx, Ly, Lz, dx):
    """threed.py:34-35."""
    nx, ny, nz = int(round(Lx / dx)), int(round(Ly / dx)), int(round(Lz / dx))
    return np.arange(nx) * dx, np.arange(ny) * dx, np.arange(nz) * dx


def trench_mask(Lx, Ly, Lz, dx, trench_width, mask_th, sub_top, hole=False):
    """Mask material nodes, threed.py:36-43 (float64, same expression order)."""
    xs, ys, zs = grid_axes(Lx, Ly, Lz, dx)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing='ij')
    mask_band = (Z >= sub_top) & (Z < sub_top + mask_th)
    if hole:
        opening = (X - Lx / 2) ** 2 + (Y - Ly / 2) ** 2 < (trench_width / 2) ** 2
    else:
        opening = np.abs(X - Lx / 2) < trench_width / 2
    return mask_band & (~opening)


def init_sdf(Lx, Ly, Lz, dx, trench_width, mask_th, sub_top, hole=False):
    """Exact signed distance to the CSG solid of threed.py:51-55 (substrate half-space union the
    masked slab), positive in solid. The reference cleans the analytic field with skfmm
    (threed.py:56); the exact distance is the quantity skfmm approximates to O(dx)."""
    xs, ys, zs = grid_axes(Lx, Ly, Lz, dx)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing='ij')
    r = np.sqrt((X - Lx / 2) ** 2 + (Y - Ly / 2) ** 2) if hole else np.abs(X - Lx / 2)
    hw, s, t = trench_width / 2.0, sub_top, sub_top + mask_th
    dr = r - hw
    d1 = np.where(r <= hw, np.abs(Z - s), np.hypot(dr, Z - s))            # opening floor
    d2 = np.where(Z < s, np.hypot(dr, Z - s),
                  np.where(Z > t, np.hypot(dr, Z - t), np.abs(dr)))        # mask sidewall
    d3 = np.where(r >= hw, np.abs(Z - t), np.hypot(dr, Z - t))            # mask top
    d = np.minimum(d1, np.minimum(d2, d3))
    # the junction plane z = sub_top under the mask is interior (substrate meets mask): the
    # analytic CSG field is 0 there, the distance is not
    solid = (Z < s) | ((Z >= s) & (Z < t) & (r > hw))
    return np.where(solid, d, -d)


# ----------------------------------------------------------------------------- advection
def advect_3d(phi, Fspeed, dx, dt):
    """threed.py:1248-1259: phi - dt*F*sqrt(sum_ax max(D-,0)^2 + min(D+,0)^2); the missing
    one-sided difference at a domain face is 0."""
    phi = np.asarray(phi, np.float64)
    g = np.zeros_like(phi)
    for ax in range(3):
        dm = np.zeros_like(phi)
        dp = np.zeros_like(phi)
        hi = [slice(None)] * 3
        lo = [slice(None)] * 3
        hi[ax] = slice(1, None)
        lo[ax] = slice(0, -1)
        diff = (phi[tuple(hi)] - phi[tuple(lo)]) / dx
        dm[tuple(hi)] = diff
        dp[tuple(lo)] = diff
        g += np.maximum(dm, 0.0) ** 2 + np.minimum(dp, 0.0) ** 2
    return phi - dt * np.asarray(Fspeed, np.float64) * np.sqrt(g)


def advect_3d_adjoint(phi, Fspeed, mask, dx, dt, gout):
    """Adjoint of one pinned advection substep (advect_3d, then out = mask_phi on mask nodes)
    at the state phi: returns (dL/dphi, dL/dF) for dL/dout = gout. With G = sum_ax
    max(D-,0)^2 + min(D+,0)^2, out = phi - dt F sqrt(G) away from the mask, so
    dout/dF = -dt sqrt(G) and dout/dphi goes through phi itself and through the one-sided
    differences of the six neighbours; where G = 0 the sqrt term has no derivative (0 is taken),
    and mask nodes pass no gradient (they are re-pinned to a constant). Not in the reference,
    which differentiates nothing in the etch loop; pinned by central differences of advect_3d."""
    phi = np.asarray(phi, np.float64)
    F = np.asarray(Fspeed, np.float64)
    free = ~np.asarray(mask, bool)
    go = np.where(free, np.asarray(gout, np.float64), 0.0)
    dmp = []
    G = np.zeros_like(phi)
    for ax in range(3):
        dm = np.zeros_like(phi)
        dp = np.zeros_like(phi)
        hi = [slice(None)] * 3
        lo = [slice(None)] * 3
        hi[ax] = slice(1, None)
        lo[ax] = slice(0, -1)
        diff = (phi[tuple(hi)] - phi[tuple(lo)]) / dx
        dm[tuple(hi)] = diff
        dp[tuple(lo)] = diff
        dmp.append((np.maximum(dm, 0.0), np.minimum(dp, 0.0), tuple(hi), tuple(lo)))
        G += np.maximum(dm, 0.0) ** 2 + np.minimum(dp, 0.0) ** 2
    sq = np.sqrt(G)
    # dL/dG at every node: gout * (-dt F / (2 sqrt G)), 0 where G = 0
    with np.errstate(divide="ignore", invalid="ignore"):
        w = np.where(sq > 0.0, -dt * F * go / (2.0 * sq), 0.0)
    gphi = go.copy()
    for mp, mn, hi, lo in dmp:
        # G_n holds max(D-_n,0)^2 with D-_n = (phi_n - phi_{n-1})/dx and min(D+_n,0)^2 with
        # D+_n = (phi_{n+1} - phi_n)/dx; scatter dG_n/dphi onto phi_n and its neighbours
        a = w * 2.0 * mp / dx                      # d/dphi_n of the D- term, -a onto phi_{n-1}
        b = w * 2.0 * mn / dx                      # d/dphi_{n+1} of the D+ term, -b onto phi_n
        gphi += a
        gphi[lo] -= a[hi]
        gphi -= b
        gphi[hi] += b[lo]
    gF = -dt * np.sqrt(G) * go
    return gphi, gF


def advect_pinned(phi, Fspeed, mask, mask_phi, dx, dt, nsub):
    """nsub substeps, each followed by the mask re-pin (threed.py:1761-1763; the device kernel
    threed.py:1268-1270 pins in the same substep)."""
    phi = np.asarray(phi, np.float64).copy()
    mask = np.asarray(mask, bool)
    mp = np.asarray(mask_phi, np.float64)
    for _ in range(nsub):
        phi = advect_3d(phi, Fspeed, dx, dt)
        phi[mask] = mp[mask]
    return phi


# ----------------------------------------------------------------------------- FSM reinit
def eik_solve(b1, b2, b3, dx):
    """threed.py:1421-1452, vectorised: sort the three upwind minima, then the 1-, 2- and
    3-term Godunov roots, taking the first whose value does not exceed the next minimum."""
    a = np.sort(np.stack([np.asarray(b1, np.float64), np.asarray(b2, np.float64),
                          np.asarray(b3, np.float64)]), axis=0)
    a1, a2, a3 = a[0], a[1], a[2]
    x1 = a1 + dx
    d2 = np.maximum(2.0 * dx * dx - (a1 - a2) ** 2, 0.0)
    x2 = 0.5 * (a1 + a2 + np.sqrt(d2))
    s1 = a1 + a2 + a3
    s2 = a1 * a1 + a2 * a2 + a3 * a3
    d3 = np.maximum(s1 * s1 - 3.0 * (s2 - dx * dx), 0.0)
    x3 = (s1 + np.sqrt(d3)) / 3.0
    return np.where(x1 <= a2, x1, np.where(x2 <= a3, x2, x3))


def _nbr(U, ax, side, fill):
    """Neighbour along `ax` (side -1 / +1) with `fill` where it falls outside the grid."""
    out = np.full_like(U, fill)
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    if side < 0:
        dst[ax] = slice(1, None)
        src[ax] = slice(0, -1)
    else:
        dst[ax] = slice(0, -1)
        src[ax] = slice(1, None)
    out[tuple(dst)] = U[tuple(src)]
    return out


def _clamped(U, ax, side):
    """Neighbour with the index clamped at the grid edge (threed.py:1463-1465)."""
    out = U.copy()
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    if side < 0:
        dst[ax] = slice(1, None)
        src[ax] = slice(0, -1)
    else:
        dst[ax] = slice(0, -1)
        src[ax] = slice(1, None)
    out[tuple(dst)] = U[tuple(src)]
    return out


def eik_init(phi0, dx, big):
    """threed.py:1455-1479: frozen interface nodes (sign change to a face neighbour) get
    min(|phi0|/max(|grad phi0|_central, 0.5), 1.8dx); the rest start at `big`."""
    c = np.asarray(phi0, np.float64)
    o = [(_clamped(c, ax, -1), _clamped(c, ax, +1)) for ax in range(3)]
    iface = np.zeros(c.shape, bool)
    for m, p in o:
        iface |= (c * m < 0.0) | (c * p < 0.0)
    g = [(p - m) * 0.5 / dx for m, p in o]
    grad0 = np.maximum(np.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2), 0.5)
    D = np.minimum(np.abs(c) / grad0, 1.8 * dx)
    U = np.where(iface, D, big)
    return U, iface


def eik_jacobi(U, frozen, dx, big):
    """threed.py:1482-1512: U <- min(U, eik_solve(min-neighbour per axis)); frozen held;
    absent neighbours at the grid edge count as `big`."""
    a = [np.minimum(_nbr(U, ax, -1, big), _nbr(U, ax, +1, big)) for ax in range(3)]
    cand = eik_solve(a[0], a[1], a[2], dx)
    return np.where(frozen, U, np.minimum(U, cand))


def reinit_fsm(phi, dx, band, n_iter=None):
    """threed.py:1515-1537."""
    phi = np.asarray(phi, np.float64)
    big = float(band + 4.0 * dx)
    if n_iter is None:
        n_iter = int(np.ceil(band / dx)) + 4
    U, frozen = eik_init(phi, dx, big)
    for _ in range(n_iter):
        U = eik_jacobi(U, frozen, dx, big)
    sgn = np.sign(phi)
    sgn = np.where(sgn == 0, 1.0, sgn)
    return sgn * np.minimum(U, band + dx)


def _eik_weights(a, U, dx):
    """Godunov root x of sum_d (x - a_d)_+^2 = dx^2 at the three axis minima a (3, ...) and its
    derivative in each: dx/da_d = (x - a_d) / sum_active (x - a_e) on the active terms (the
    implicit function theorem on the active equation), 0 on the others."""
    x = eik_solve(a[0], a[1], a[2], dx)
    act = a < x[None]
    r = np.where(act, x[None] - a, 0.0)
    den = r.sum(axis=0)
    with np.errstate(divide="ignore", invalid="ignore"):
        w = np.where(den[None] > 0.0, r / den[None], 0.0)
    return x, w


def reinit_fsm_adjoint(phi, dx, band, gout, n_iter=None, state_dtype=np.float64):
    """d loss / d phi of reinit_fsm(phi, dx, band) for d loss / d out = gout: the exact adjoint
    of the piecewise-smooth map the sweeps compute, with every discrete choice (interface flags,
    the smaller neighbour per axis, the active Godunov terms, whether a sweep lowers a node,
    the clamp at band + dx) held at the forward's. Each sweep's state is recomputed and the
    sweeps are transposed in reverse: a node lowered by sweep k took eik_solve of its upwind
    neighbours' values at k, whose derivative is the Godunov weight; the frozen seeds
    |phi|/max(|grad phi|_central, 0.5) pass their derivative to the node and its six
    neighbours. Not in the reference, which differentiates nothing in its etch loop; pinned by
    central differences of reinit_fsm.

    state_dtype=np.float32 rounds every sweep state (and each Godunov root) to float32 before
    the next sweep reads it, as a device that stores the states in float32 does; its decisions
    (which sweep lowers a node) then follow that rounding."""
    c = np.asarray(phi, np.float64)
    big = float(band + 4.0 * dx)
    if n_iter is None:
        n_iter = int(np.ceil(band / dx)) + 4
    U, frozen = eik_init(c, dx, big)
    U = U.astype(state_dtype).astype(np.float64)
    states = [U]
    for _ in range(n_iter):
        U = eik_jacobi(U, frozen, dx, big).astype(state_dtype).astype(np.float64)
        states.append(U)
    sgn = np.where(np.sign(c) == 0, 1.0, np.sign(c))
    G = np.where(states[-1] < band + dx, np.asarray(gout, np.float64) * sgn, 0.0)
    for k in range(n_iter - 1, -1, -1):
        Uk = states[k]
        lo = [_nbr(Uk, ax, -1, big) for ax in range(3)]
        hi = [_nbr(Uk, ax, +1, big) for ax in range(3)]
        a = np.stack([np.minimum(lo[ax], hi[ax]) for ax in range(3)])
        x, w = _eik_weights(a, Uk, dx)
        took = ~frozen & (x.astype(state_dtype).astype(np.float64) < Uk)
        Gt = np.where(took, G, 0.0)
        Gn = np.where(took, 0.0, G)
        for ax in range(3):
            # the smaller neighbour (the lower side on a tie) receives w * G
            use_lo = lo[ax] <= hi[ax]
            contrib = w[ax] * Gt
            to_lo = np.where(use_lo, contrib, 0.0)
            to_hi = np.where(use_lo, 0.0, contrib)
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            src[ax], dst[ax] = slice(1, None), slice(0, -1)     # node n sends to n - 1
            Gn[tuple(dst)] += to_lo[tuple(src)]
            src[ax], dst[ax] = slice(0, -1), slice(1, None)     # node n sends to n + 1
            Gn[tuple(dst)] += to_hi[tuple(src)]
        G = Gn
    # the frozen seeds D = min(|c| / max(|g|, 0.5), 1.8 dx), g the central gradient with the
    # index clamped at the grid edge (a clamped neighbour is the node itself)
    o = [(_clamped(c, ax, -1), _clamped(c, ax, +1)) for ax in range(3)]
    g = [(p - m) * 0.5 / dx for m, p in o]
    gn = np.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
    grad0 = np.maximum(gn, 0.5)
    D = np.abs(c) / grad0
    live = frozen & (D < 1.8 * dx)
    gD = np.where(live, G, 0.0)
    gc = gD * np.sign(c) / grad0
    # through |g|: dD/dg_ax = -|c| g_ax / (grad0^2 |g|) where |g| > 0.5
    steep = live & (gn > 0.5)
    with np.errstate(divide="ignore", invalid="ignore"):
        coef = np.where(steep, -gD * np.abs(c) / (grad0 * grad0 * gn), 0.0)
    for ax in range(3):
        t = coef * g[ax] * 0.5 / dx          # d/dc_{n+1}: +t, d/dc_{n-1}: -t
        n_ax = c.shape[ax]
        idx = np.arange(n_ax)
        shp = [1, 1, 1]
        shp[ax] = n_ax
        at_hi = (idx == n_ax - 1).reshape(shp)
        at_lo = (idx == 0).reshape(shp)
        # the +1 neighbour (the node itself at the upper edge)
        gc += np.where(at_hi, t, 0.0)
        src = [slice(None)] * 3
        dst = [slice(None)] * 3
        src[ax], dst[ax] = slice(0, -1), slice(1, None)
        gc[tuple(dst)] += np.where(at_hi, 0.0, t)[tuple(src)]
        # the -1 neighbour (the node itself at the lower edge)
        gc -= np.where(at_lo, t, 0.0)
        src[ax], dst[ax] = slice(1, None), slice(0, -1)
        gc[tuple(dst)] -= np.where(at_lo, 0.0, t)[tuple(src)]
    return gc


# ----------------------------------------------------------------------------- depth metrics
def _floor_z(col, zs):
    """Lowest z of the gas run that reaches the top (threed.py:1824-1831)."""
    k = len(col) - 1
    if not col[k]:
        return None
    while k > 0 and col[k - 1]:
        k -= 1
    return zs[k]


def depth3d(geo, half=1.0):
    """threed.py:1798-1834 (center depth: median floor over central columns)."""
    phi, xs, ys, zs = geo['phi'], geo['xs'], geo['ys'], geo['zs']
    cx, cy = geo['Lx'] / 2, geo['Ly'] / 2
    r_feat = geo.get('trench_width', 2.0) / 2.0 - geo['dx']
    hx = max(min(half, r_feat), geo['dx'])
    ic = np.where(np.abs(xs - cx) <= hx)[0]
    if geo.get('hole', False):
        jc = np.where(np.abs(ys - cy) <= hx)[0]
        inside = lambda i, j: (xs[i] - cx) ** 2 + (ys[j] - cy) ** 2 <= r_feat ** 2
    else:
        jc = np.where(np.abs(ys - cy) <= half)[0]
        inside = lambda i, j: True
    floors = []
    for i in ic:
        for j in jc:
            if not inside(i, j):
                continue
            z = _floor_z(phi[i, j, :] < 0, zs)
            if z is not None:
                floors.append(z)
    if not floors:
        return 0.0
    return float(geo['sub_top'] - np.median(floors))


def max_depth_3d(geo):
    """threed.py:1841-1869 (5th-percentile floor over the feature footprint)."""
    phi, xs, ys, zs = geo['phi'], geo['xs'], geo['ys'], geo['zs']
    cx, cy = geo['Lx'] / 2.0, geo['Ly'] / 2.0
    r = geo.get('trench_width', 4.0) / 2.0
    if geo.get('hole', False):
        ii, jj = np.where((xs[:, None] - cx) ** 2 + (ys[None, :] - cy) ** 2 <= (r + geo['dx']) ** 2)
        cols = list(zip(ii, jj))
    else:
        ix = np.where(np.abs(xs - cx) <= r + geo['dx'])[0]
        cols = [(i, j) for i in ix for j in range(len(ys))]
    floors = []
    for i, j in cols:
        z = _floor_z(phi[i, j, :] < 0, zs)
        if z is not None:
            floors.append(z)
    if not floors:
        return 0.0
    return float(geo['sub_top'] - np.percentile(floors, 5))


# ----------------------------------------------------------------------------- profile r(z)
def smooth_heaviside(x, eps):
    """0 below -eps, 1 above eps, (1 + x/eps + sin(pi x/eps)/pi)/2 between (C1-smooth)."""
    x = np.asarray(x, np.float64)
    mid = 0.5 * (1.0 + x / eps + np.sin(np.pi * x / eps) / np.pi)
    return np.where(x <= -eps, 0.0, np.where(x >= eps, 1.0, mid))


def slice_gas_area(phi, dx, eps):
    """A(z) = dx^2 * sum over each z slice of H_eps(-phi): the smoothed gas cross-section area.
    Not in the reference (whose depth metrics are whole-cell column floors); the profile the
    differentiable configuration's loss is written on, r(z) = sqrt(A(z) / pi) for a hole."""
    return smooth_heaviside(-np.asarray(phi, np.float64), eps).sum(axis=(0, 1)) * dx * dx


def slice_gas_area_adjoint(phi, dx, eps, garea):
    """d loss / d phi of slice_gas_area for garea = d loss / dA (one value per z slice)."""
    x = -np.asarray(phi, np.float64)
    dH = np.where(np.abs(x) < eps, 0.5 * (1.0 + np.cos(np.pi * x / eps)) / eps, 0.0)
    return -np.asarray(garea, np.float64)[None, None, :] * dx * dx * dH


def profile_rz(phi, dx, eps):
    """Equivalent radius sqrt(A(z) / pi) of every z slice."""
    return np.sqrt(np.maximum(slice_gas_area(phi, dx, eps), 0.0) / np.pi)


def gaussian_smooth(phi, sigma, truncate=4.0):
    """scipy.ndimage.gaussian_filter(phi, sigma) as run_etch_3d's surf_smooth calls it
    (threed.py:1779-1782): weights exp(-x^2 / (2 sigma^2)) over |x| <= int(truncate sigma + 0.5),
    normalised; one axis after the other (x, y, z), each pass in float64 and stored in the
    input's dtype; boundaries 'reflect' (d c b a | a b c d | d c b a); the symmetric taps summed
    as w0 x0 + sum over j = r..1 of wj (x[-j] + x[+j]). Pinned bit for bit against scipy
    (tests/test_oracle_levelset.py)."""
    a = np.array(phi)
    r = int(truncate * float(sigma) + 0.5)
    x = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 / (float(sigma) ** 2) * x ** 2)
    w = w / w.sum()
    for ax in range(a.ndim):
        n = a.shape[ax]
        src = np.moveaxis(a, ax, -1).astype(np.float64)

        def at(k):
            k = np.arange(n) + k
            while True:                            # reflect until every index is inside
                lo, hi = k < 0, k >= n
                if not (lo.any() or hi.any()):
                    return src[..., k]
                k = np.where(lo, -k - 1, np.where(hi, 2 * n - k - 1, k))
        acc = w[r] * src
        for j in range(r, 0, -1):
            acc = acc + (at(-j) + at(j)) * w[r + j]
        a = np.moveaxis(acc.astype(a.dtype), -1, ax)
    return a


def reinit_pde(phi, dx, n_iter=24):
    """reinit_gpu (threed.py:1349-1418): n_iter forward-Euler steps (dtau = 0.3 dx) of
    phi_tau + sgn(phi0)(|grad phi| - 1) = 0 in float32, one operation one rounding, neighbours
    clamped at the boundary. Interface nodes (a sign change of phi0 to a face neighbour) relax
    toward D = clip(phi0 / max(|grad phi0|_central, 0.5), +-1.8 dx) (the Russo-Smereka fix);
    the others take the Godunov upwind |grad phi| with the smoothed sign
    phi0 / sqrt(phi0^2 + (dx/2)^2)."""
    f = np.float32
    p0 = np.asarray(phi, np.float32)
    a = p0.copy()
    inv = f(1.0) / f(dx)
    dtau = f(f(0.3) * f(dx))
    nx, ny, nz = p0.shape

    def sh(x, ax, d):                          # x at index i + d along ax, clamped
        idx = np.clip(np.arange(x.shape[ax]) + d, 0, x.shape[ax] - 1)
        return np.take(x, idx, axis=ax)
    o = [sh(p0, ax, d) for ax in range(3) for d in (-1, 1)]
    iface = np.zeros(p0.shape, bool)
    for q in o:
        iface |= (p0 * q) < f(0.0)
    half = f(0.5)
    g0 = [((o[2 * ax + 1] - o[2 * ax]) * half) * inv for ax in range(3)]
    grad0 = np.maximum(np.sqrt(g0[0] * g0[0] + g0[1] * g0[1] + g0[2] * g0[2]), f(0.5))
    dxl = f(1.0) / inv
    lim = f(f(1.8) * dxl)
    D = np.clip(p0 / grad0, -lim, lim)
    sgn0 = np.where(p0 < f(0.0), f(-1.0), f(1.0))
    sgn = p0 / np.sqrt(p0 * p0 + f(0.25) / (inv * inv))
    for _ in range(int(n_iter)):
        c = a
        out_if = c - (dtau * inv) * (sgn0 * np.abs(c) - D)
        Dm = [(c - sh(c, ax, -1)) * inv for ax in range(3)]
        Dp = [(sh(c, ax, 1) - c) * inv for ax in range(3)]
        z = f(0.0)
        gp = [np.maximum(np.maximum(Dm[ax], z), -np.minimum(Dp[ax], z)) for ax in range(3)]
        gn = [np.maximum(-np.minimum(Dm[ax], z), np.maximum(Dp[ax], z)) for ax in range(3)]
        g = [np.where(sgn > z, gp[ax], gn[ax]) for ax in range(3)]
        G = np.sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2])
        out = c - dtau * sgn * (G - f(1.0))
        a = np.where(iface, out_if, out).astype(np.float32)
    return a
