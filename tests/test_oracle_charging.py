"""Pins of the oracle's surface-charging hooks (oracle/charging.py): the tables at their nodes
and ends, the alpha factor's limits, the throttle's floor scaling, and the Hwang-Giapis
closure's yield ratio (closed form in the mean-energy mode) and conservation of the deflected
ions."""
import numpy as np
import pytest

from oracle import charging as och
from petch.params import PAR


def test_tables_at_nodes_between_and_beyond():
    assert och.floor_profile(2.0) == pytest.approx((0.40, 19.0, 20.0))
    assert och.floor_profile(0.0) == pytest.approx((1.0, 0.0, 0.0))
    assert och.floor_profile(1.1) == pytest.approx((0.57, 9.0, 15.75))    # midway 1.0-1.2
    assert och.floor_profile(9.0) == pytest.approx((0.22, 33.0, 27.5))    # held at the end
    assert och.floor_profile(0.5, "petch") == pytest.approx((0.704, 9.5, 16.6))
    assert och.floor_profile(3.0, "petch") == pytest.approx((0.406, 27.7, 18.8))
    assert och.throttle_q(-1.0) == 1.0 and och.throttle_q(0.0) == 1.0
    assert och.throttle_q(8.0) == pytest.approx(0.25)
    assert och.throttle_q(11.5) == pytest.approx(0.425)                  # midway 8-15
    assert och.throttle_q(40.0) == pytest.approx(0.5)


def test_alpha_factor_limits():
    rng = np.random.default_rng(0)
    m = rng.uniform(0.01, 2.0, 500)
    np.testing.assert_allclose(och.charge_factor(m, m, 0.7), 1.0)       # same shadowing
    f = och.charge_factor(m, np.zeros_like(m), 0.7)                      # electrons all shadowed
    np.testing.assert_allclose(f, 0.3)
    f = och.charge_factor(m, 2 * m, 0.7)                                 # electrons less shadowed
    assert (f <= 1.0).all() and (f >= 0.3).all()


def _faces():
    rng = np.random.default_rng(1)
    n = 400
    c = np.stack([rng.uniform(0, 10, n), rng.uniform(0, 4, n), rng.uniform(2, 14, n)], 1)
    gz = rng.uniform(-1, 1, n)
    geo = dict(Lx=10.0, Ly=4.0, dx=0.25, trench_width=2.0, sub_top=12.0, hole=False)
    return c, gz, geo, rng


def test_throttle_scales_the_floor_only():
    c, gz, geo, rng = _faces()
    m = rng.uniform(0.1, 1.0, len(c))
    iy = tuple(rng.uniform(0.1, 1.0, len(c)) for _ in range(3))
    _, _, floor = och.masks(c, gz, geo)
    assert floor.any() and (~floor).any()
    Q = och.throttle_q((geo['sub_top'] - c[floor, 2].min()) / geo['trench_width'])
    m2, iy2 = och.floor_throttle(c, gz, geo, m, iy)
    np.testing.assert_allclose(m2, np.where(floor, Q * m, m))
    for a, b in zip(iy, iy2):
        np.testing.assert_allclose(b, np.where(floor, Q * a, a))


def test_hg_closure_yield_ratio_and_conservation():
    """Mean-energy yields A (sqrt(E) - sqrt(Eth)): the floor's ion-enhanced yield scales by
    Q (sqrt(E - Vf) - sqrt(Eth_ie)) / (sqrt(E) - sqrt(Eth_ie)); the foot receives, summed over
    its area, exactly the area-weighted ions taken off the floor times A_ie (sqrt(E_defl) -
    sqrt(Eth_ie))."""
    c, gz, geo, rng = _faces()
    par = dict(PAR, ied_mode='mean', Emean=120.0)
    A = rng.uniform(0.01, 0.05, len(c))
    m = rng.uniform(0.1, 1.0, len(c))
    iy = tuple(rng.uniform(0.1, 1.0, len(c)) for _ in range(3))
    inside, below, floor = och.masks(c, gz, geo)
    zf = c[floor, 2].min()
    Q, Vf, E = och.floor_profile((geo['sub_top'] - zf) / geo['trench_width'])
    m2, (sp2, ie2, ip2) = och.hg_closure(c, gz, A, geo, par, m, iy)
    s = np.sqrt
    r_ie = (s(120.0 - Vf) - s(par['Eth_ie'])) / (s(120.0) - s(par['Eth_ie']))
    np.testing.assert_allclose(ie2[floor], iy[1][floor] * Q * r_ie, rtol=1e-12)
    np.testing.assert_allclose(m2[floor], Q * m[floor], rtol=1e-12)
    foot = inside & below & (gz <= 0.7) & (c[:, 2] <= zf + max(0.3 * 2.0, 0.5))
    assert foot.any()
    moved = ((ie2 - iy[1]) * A)[foot].sum()
    expect = (m[floor] * A[floor]).sum() * (1 - Q) * par['A_ie'] * max(s(E) - s(par['Eth_ie']), 0)
    assert expect > 0 and moved == pytest.approx(expect, rel=1e-10)
    other = ~(floor | foot)
    np.testing.assert_array_equal(ie2[other], iy[1][other])
