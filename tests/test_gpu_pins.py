"""The HIP traces against closed forms that do not go through the oracle: the RR neutral in a
closed cavity (E[deposit per ray] = 1 / S), the faithful ion reflected off a steep wall onto a
floor (the published sticking ramp and reflected-energy law by quadrature), and the RR mean in
a deep hole against the radiosity solution of the same surface, built from the HIP form-factor
rays and solved with scipy. Each case also checks the HIP trace against the oracle's per-face
sums on the same streams.

The per-face accumulators are float32 sums of float atomic adds, as in the reference
(threed.py:579,650-654); a sum of k adds carries a relative rounding error of up to about
k * 2^-24 (ulp(sum) / 2 per add), and with hundreds of thousands of adds on one face that
reaches 1e-3 and more. The meshes and ray counts here keep each face to a few thousand adds, and
the bars include that rounding bound."""
import numpy as np
import pytest
from scipy import integrate, stats

from pin_util import closed_box, reflected_energy_expectation, rr_hit_law, wall_and_floor

from oracle import chem as oc
from oracle import levelset as ol
from oracle import native as on
from petch.params import PAR

pytestmark = pytest.mark.gpu


def _dev():
    import torch
    return torch.device("cuda", 0)


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(_dev())


def _rel_l1(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-30)


@pytest.mark.parametrize("S", [0.05, 0.3, 0.7])
def test_rr_closed_cavity_unbiased(S):
    from petch import ops
    Lx, Ly = 2.0, 1.5
    v, f = closed_box((-0.3, -0.3, 0.0), (Lx + 0.3, Ly + 0.3, 2.8), n=24)
    bvh = ops.Bvh(_dev()).build(_t(v), _t(f))
    box = ops.make_box(Lx, Ly, 3.0, 2.5, 0)
    beta = 0.5
    bare = np.full(len(f), S / beta, np.float32)
    S32 = float(np.float32(S / beta) * np.float32(beta))     # the sticking the trace applies
    n = 200_000
    print(f"closed cavity S={S}: {len(f)} faces, {n} rays", flush=True)
    fl = ops.trace_neutral_rr(bvh, box, ops.make_source(0), _t(bare), beta, n, 77).cpu().numpy()
    assert bvh.status() == 0
    m1, c, p1, p2, T0, Tc = rr_hit_law(S32)
    EJ = p1 / (1 - p2)
    EJ2 = p1 * (1 + p2) / (1 - p2) ** 2
    mean = T0 + Tc * EJ
    se = np.sqrt(Tc ** 2 * (EJ2 - EJ ** 2) / n)
    hits = n * (m1 + c * EJ)                                # expected adds over all faces
    rounding = hits / len(f) * 2.0 ** -24 * mean            # per-ray bound of the float32 sums
    got = fl.astype(np.float64).sum() / n
    print(f"  deposit per ray {got:.6f} closed form {mean:.6f} = 1/S {1 / S32:.6f} (se {se:.1e}, "
          f"float32 rounding bound {rounding:.1e})", flush=True)
    assert mean == pytest.approx(1 / S32, rel=1e-9)
    assert abs(got - mean) < 4 * se + rounding
    # the same streams as the oracle (fewer rays: about 50 adds per face). A path that meets a
    # face edge within the float32 rounding of its hit point splits from the oracle's (double)
    # one, and in a closed box every later bounce of it differs: with 3-4 bounces per ray
    # (S = 0.7, 0.3) that stays below 1e-4 of the deposits; with ~55 (S = 0.05) whole paths
    # decorrelate, and only the closed-form total above compares
    if S >= 0.3:
        ob = on.OracleBvh(v, f).trace_neutral_rr(on.Box(Lx, Ly, 3.0, 2.5, 0),
                                                 on.Source(0, 0, 0), bare, beta, 20000, 77)
        gs = ops.trace_neutral_rr(bvh, box, ops.make_source(0), _t(bare), beta, 20000,
                                  77).cpu().numpy()
        r = _rel_l1(gs, ob)
        print(f"  same-seed L1 vs oracle {r:.2e}", flush=True)
        assert r < 1e-4


@pytest.mark.parametrize("deg", [75.0, 80.0, 85.0])
def test_ion_grazing_reflection_closed_form(deg):
    from petch import ops
    theta = np.deg2rad(deg)
    Lx, Ly = 1.0, 1.0
    v, f, n_wall, H = wall_and_floor(theta, Lx, Ly, n=48)
    bvh = ops.Bvh(_dev()).build(_t(v), _t(f))
    box = ops.make_box(Lx, Ly, H + 1.0, H + 0.5, 0)
    ip = oc.ion_params(dict(PAR))
    n = 1_000_000
    print(f"grazing ion {deg} deg: {n} ions", flush=True)
    out = ops.trace_ion_yield(bvh, box, ops.make_source(1, 0.0), ops.IonParams(**ip), n,
                              101).cpu().numpy().astype(np.float64)
    assert bvh.status() == 0
    wall, floor = out[:, :n_wall].sum(1) / n, out[:, n_wall:].sum(1) / n
    c = np.cos(theta)
    f_sp = (1 + ip['B_sp'] * (1 - c * c)) * c
    f_ie = max(3 - 6 * theta / np.pi, 0.0) if c < 0.5 else 1.0
    stick = 1 - np.clip((theta - ip['thrmin']) / (ip['thrmax'] - ip['thrmin']), 0, 1)
    if theta >= ip['inflect']:
        Eref = 1 - (1 - ip['A_en']) * (np.pi / 2 - theta) / (np.pi / 2 - ip['inflect'])
    else:
        Eref = ip['A_en'] * (theta / ip['inflect']) ** ip['n_l']
    pdfE = lambda E: stats.norm.pdf(E, ip['meanE'], ip['sigmaE'])
    lim = (ip['meanE'] - 8 * ip['sigmaE'], ip['meanE'] + 8 * ip['sigmaE'])

    def yield_of(A, Eth):
        return lambda x: A * max(np.sqrt(max(x, 0.0)) - np.sqrt(Eth), 0.0)

    def over_E(fn):
        return integrate.quad(lambda E: fn(E) * pdfE(E), *lim, limit=200)[0]

    # every ion hits the wall once: n / n_wall float32 adds per wall face
    rounding = n / n_wall * 2.0 ** -24
    assert wall[3] == pytest.approx(1.0, rel=rounding)
    assert wall[4] == pytest.approx(c, rel=1e-5 + rounding)
    for ch, (A, Eth, fa) in enumerate([(ip['A_sp'], ip['Eth_sp'], f_sp),
                                      (ip['A_ie'], ip['Eth_ie'], f_ie),
                                      (ip['A_p'], ip['Eth_p'], f_ie)]):
        assert wall[ch] == pytest.approx(over_E(lambda E: yield_of(A, Eth)(E) * fa), rel=1e-3)
    ws = 1 - stick
    arr = ws * over_E(lambda E: reflected_energy_expectation(lambda x: 1.0, E, Eref, ip['minE']))
    ie = ws * over_E(lambda E: reflected_energy_expectation(yield_of(ip['A_ie'], ip['Eth_ie']),
                                                           E, Eref, ip['minE']))
    print(f"  floor arr {floor[3]:.5f} / {arr:.5f}, ie {floor[1]:.5f} / {ie:.5f}", flush=True)
    assert floor[3] == pytest.approx(arr, rel=2e-3, abs=1e-4)
    assert floor[1] == pytest.approx(ie, rel=2e-3)
    ob = on.OracleBvh(v, f).trace_ion_yield(on.Box(Lx, Ly, H + 1.0, H + 0.5, 0),
                                            on.Source(1, 0.0, 0.0), on.Ion(**ip), 200000, 101)
    gs = ops.trace_ion_yield(bvh, box, ops.make_source(1, 0.0), ops.IonParams(**ip), 200000,
                             101).cpu().numpy()
    # the reflected ions' lobes come from sines and cosines whose last bits differ between the
    # device and libm: a few floor hits move across a face edge (measured up to 1.6e-5)
    for ch in range(5):
        r = _rel_l1(gs[ch], ob[ch])
        print(f"  channel {ch} same-seed L1 vs oracle {r:.2e}", flush=True)
        assert r < 1e-4


def test_rr_mean_matches_radiosity_of_the_hip_form_factors():
    """The HIP RR trace in a deep hole (16M walks) against G_i = D_i + sum_j (A_j / A_i) F_ji
    (1 - S_j) G_j solved by scipy from the HIP form-factor rays (1024 per face): z-band averages
    within 2 % (tests/test_oracle_pins.py has the same check on the oracle)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    from petch import flux, ops
    from oracle import flux as ofl
    Lx = Ly = 8.0
    Lz, dx = 7.0, 0.125
    w, mask_th, sub = 1.5, 5.0, 1.5
    phi = ol.init_sdf(Lx, Ly, Lz, dx, w, mask_th, sub, True).astype(np.float32)
    tphi = _t(phi)
    mesh = flux.SurfaceMesh(_dev()).build(tphi, dx)
    nf = mesh.nf
    box = ops.make_box(Lx, Ly, Lz, Lz - dx, 0)
    bare = np.random.default_rng(5).uniform(0.2, 1.0, nf).astype(np.float32)
    beta = 0.5
    S = bare.astype(np.float64) * beta
    n = 16_000_000
    print(f"deep hole: {nf} faces, {n} walks", flush=True)
    fl = ops.trace_neutral_rr(mesh.bvh, box, ops.make_source(0), _t(bare), beta, n,
                              41).cpu().numpy().astype(np.float64)
    A = np.maximum(mesh.areas.cpu().numpy().astype(np.float64), 1e-12)
    m_mc = fl / A * (Lx * Ly / n)
    gn = ops.gas_normals(tphi, dx, mesh.centroids, mesh.normals)
    n_ff = 1024
    hit = ops.trace_ff(mesh.bvh, box, mesh.centroids, gn, n_ff, 43).cpu().numpy()
    assert mesh.bvh.status() == 0
    D, Aff = ofl.radiosity_matrix(hit, nf, n_ff)
    Fa = (sp.diags(1.0 / A) @ Aff @ sp.diags(A)).multiply((1.0 - S)[None, :]).tocsc()
    G = spl.spsolve(sp.identity(nf, format='csc') - Fa, D)
    c = mesh.centroids.cpu().numpy().astype(np.float64)
    r = np.hypot(c[:, 0] - Lx / 2, c[:, 1] - Ly / 2)
    z = c[:, 2]
    inside = (r < w / 2 + dx) & (z < sub + mask_th - dx / 2) & (A > 1e-9)
    edges = np.linspace(z[inside].min() - 1e-6, sub + mask_th, 7)
    rel = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = inside & (z >= lo) & (z < hi)
        a, b = ((A[sel] * x[sel]).sum() / A[sel].sum() for x in (m_mc, G))
        rel.append(a / b - 1)
        print(f"  z [{lo:.2f},{hi:.2f}): MC {a:.4f} radiosity {b:.4f} rel {a / b - 1:+.4f}",
              flush=True)
    assert np.max(np.abs(rel)) < 0.02, rel
