"""Geometries and closed forms shared by the oracle and GPU pin tests: plane patches, a closed
box, a steep wall over a floor, the hit law of the RR neutral in a cavity (threed.py:578-590)
and the reflected-ion energy law (threed.py:682-693)."""
import numpy as np
from scipy import integrate, stats


def grid_patch(p0, e1, e2, n1, n2):
    """Triangles of the parallelogram p0 + a e1 + b e2, a, b in [0, 1], on an n1 x n2 grid."""
    p0, e1, e2 = (np.asarray(x, np.float64) for x in (p0, e1, e2))
    a = np.linspace(0, 1, n1 + 1)
    b = np.linspace(0, 1, n2 + 1)
    A, B = np.meshgrid(a, b, indexing='ij')
    v = p0 + A.ravel()[:, None] * e1 + B.ravel()[:, None] * e2
    f = []
    for i in range(n1):
        for j in range(n2):
            q = i * (n2 + 1) + j
            f += [(q, q + n2 + 1, q + n2 + 2), (q, q + n2 + 2, q + 1)]
    return v, np.array(f, np.int64)


def join(*parts):
    vs, fs, off = [], [], 0
    for v, f in parts:
        vs.append(v)
        fs.append(f + off)
        off += len(v)
    return np.concatenate(vs).astype(np.float32), np.concatenate(fs).astype(np.int32)


def closed_box(lo, hi, n=4):
    """The six faces of the box [lo, hi] (shared corner vertices are duplicated, but every edge
    of one face lies exactly on an edge of its neighbour, so no ray passes between them)."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    ex, ey, ez = (np.diag(hi - lo)[k] for k in range(3))
    return join(grid_patch(lo, ex, ey, n, n), grid_patch(lo + ez, ex, ey, n, n),
                grid_patch(lo, ex, ez, n, n), grid_patch(lo + ey, ex, ez, n, n),
                grid_patch(lo, ey, ez, n, n), grid_patch(lo + ex, ey, ez, n, n))


def rr_hit_law(S):
    """Closed form of the RR rule of threed.py:578-590 for a ray that never escapes, uniform
    sticking S: the weights (1 - S)^j are deposited until the first one that falls below 0.1
    after its update (m1 deposits), which survives the roulette with p1 = w / 0.3; each survivor
    restarts at 0.3 and deposits c more weights 0.3 (1 - S)^j before the next roulette, survived
    with p2 = (1 - S)^c. Returns (m1, c, p1, p2, T(m1), T per cycle)."""
    q = 1.0 - S
    m1 = 1
    while q ** m1 >= 0.1:
        m1 += 1
    c = 1
    while 0.3 * q ** c >= 0.1:
        c += 1
    T0 = sum(q ** j for j in range(m1))
    Tc = sum(0.3 * q ** j for j in range(c))
    return m1, c, q ** m1 / 0.3, q ** c, T0, Tc


def wall_and_floor(theta, Lx, Ly, n=12):
    """A steep plane whose up-facing normal is theta from the vertical, covering the source
    footprint [0, Lx] x [0, Ly] (vertical ions hit it at incidence theta), and a floor at z = 0
    from its foot outwards that catches every reflected ion; n x n/2 quads on the wall, n x n on
    the floor."""
    xb = Lx + 0.05
    cot = np.cos(theta) / np.sin(theta)
    H = (Lx + 0.1) / cot
    foot = (xb, -0.6, 0.0)
    wall = grid_patch(foot, (-H * cot, 0.0, H), (0.0, Ly + 1.2, 0.0), n, n // 2)
    reach = H * np.tan(np.deg2rad(50)) + 1.0
    side = H * np.tan(np.deg2rad(20)) + 1.0
    floor = grid_patch((xb, -side, 0.0), (reach, 0.0, 0.0), (0.0, Ly + 2 * side, 0.0), n, n)
    v, f = join(wall, floor)
    return v, f, 2 * n * (n // 2), H


def reflected_energy_expectation(g, E, Eref, minE):
    """E[g(E') 1{E' > minE}] for the reflected energy of threed.py:682-693: up to four draws of
    N(Eref E, 0.1 E), the first in [0, E] taken, else Eref E."""
    mu, sd = Eref * E, 0.1 * E
    p = stats.norm.cdf(E, mu, sd) - stats.norm.cdf(0.0, mu, sd)
    h = lambda x: g(x) * (x > minE)
    acc = integrate.quad(lambda x: h(x) * stats.norm.pdf(x, mu, sd), max(minE, 0.0), E,
                         limit=200, points=[mu])[0]
    q4 = (1 - p) ** 4
    return (1 - q4) * acc / p + q4 * h(mu) if p > 0 else h(mu)
