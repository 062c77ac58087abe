"""Host-side logic of the drop-in package, checked without a GPU: parameter derivation, the CFL
substep rule, the reinit iteration count, the API surface, and that the GPU path fails loudly
instead of falling back to the CPU."""
import math

import numpy as np
import pytest

import petch
from oracle import chem as oc
from petch import flux
from petch.chemistry import _yields, angular_factors, chem_params
from petch.params import PAR, Flags


@pytest.mark.parametrize("mode", ["mean", "gauss", "bimodal"])
def test_yields_match_oracle(mode):
    par = dict(PAR)
    par['ied_mode'] = mode
    np.testing.assert_allclose(_yields(par), oc.yields(par), rtol=1e-15)
    # 'mean' is the closed form A*(sqrt(E)-sqrt(Eth))
    if mode == "mean":
        assert _yields(par)[0] == pytest.approx(7.0 * (10.0 - math.sqrt(15.0)))


def test_yield_gauss_below_mean_by_jensen():
    par = dict(PAR)
    y_mean = _yields(dict(par, ied_mode='mean'))
    y_gauss = _yields(dict(par, ied_mode='gauss'))
    assert all(g < m for g, m in zip(y_gauss, y_mean))    # sqrt is concave


def test_angular_factors_closed_forms():
    c = np.array([1.0, 0.5, 0.49, 0.0, -0.2, 1.3])
    f_ie, f_sp = angular_factors(c, PAR, "viennaps")
    assert f_ie[0] == 1.0 and f_ie[1] == 1.0 and f_ie[3] == 0.0
    assert f_ie[2] == pytest.approx(3 - 6 * math.acos(0.49) / math.pi)
    assert f_sp[0] == pytest.approx(1.0) and f_sp[3] == 0.0
    assert f_sp[1] == pytest.approx((1 + 9.3 * 0.75) * 0.5)
    f_ie, f_sp = angular_factors(c, PAR, "cosine")
    np.testing.assert_array_equal(f_ie, np.clip(c, 0, 1))


def test_chem_params_and_ion_params():
    fl = Flags(chemistry="belen", yield_angular="viennaps")
    cp = chem_params(dict(PAR), fl)
    assert cp.chemistry_belen == 1 and cp.angular_viennaps == 1
    assert cp.Yie == pytest.approx(oc.yields(PAR)[0])
    assert chem_params(dict(PAR), None).chemistry_belen == 0
    with pytest.raises(ValueError):
        chem_params(dict(PAR), Flags(chemistry="bogus"))
    ip = flux.ion_params(dict(PAR))
    ref = oc.ion_params(dict(PAR))
    for k, v in ref.items():
        assert getattr(ip, k) == pytest.approx(v, rel=1e-7), k
    # A_en of threed.py:801 for the default inflect angle
    assert ip.A_en == pytest.approx(1.0 / (1.0 + 10.0 * (1.5707963 / 1.55334303 - 1.0)), rel=1e-6)


@pytest.mark.parametrize("vmax,dt,dx,expect", [
    (0.0, 0.075, 0.25, 1),                 # floored at 1e-6 -> 1 substep
    (float("nan"), 0.075, 0.25, 1),
    (10.0, 0.075, 0.25, 8),                # ceil(0.75/0.1) = 8
    (1e6, 0.075, 0.25, 160),               # capped
    (4.0, 0.1, 0.1, 10),                   # exactly 10 -> 10
])
def test_cfl_substeps(vmax, dt, dx, expect):
    assert flux.cfl_substeps(vmax, dt, dx)[0] == expect


def test_reinit_iteration_count_quirk():
    """n_iter = ceil(band/dx) + 4 evaluated in float64 (threed.py:1526): for dx = 0.1 the driver's
    band 4dx + 2dx is 0.6000000000000001, so ceil gives 7 and n_iter = 11; for dx = 0.25, 10."""
    for dx, expect in ((0.1, 11), (0.25, 10), (0.15, 10)):
        band = 4 * dx + 2.0 * dx
        assert int(np.ceil(band / dx)) + 4 == expect


def test_api_surface_matches_reference_names():
    for name in ("Domain", "SF6O2", "Process", "Result", "PAR", "Flags", "DEFAULT_FLAGS"):
        assert hasattr(petch, name)
    from petch import threed
    for name in ("run_etch_3d", "make_trench_3d", "extract_mesh_3d", "extract_mesh_3d_gpu",
                 "mc_flux_3d_coupled", "mc_flux_3d_radiosity", "mc_flux_3d", "advect_3d_gpu",
                 "advect_3d", "reinit_fsm", "extend_velocity_gpu", "gpu_warmstart_bare",
                 "faces_in_mask", "center_depth_3d", "max_depth_3d", "smooth_flux"):
        assert callable(getattr(threed, name)), name
    dom = petch.Domain.hole(extent=14, dx=0.25, diameter=6, mask=2, depth=18)
    assert dom._geo == dict(Lx=14, Ly=14, Lz=24.0, dx=0.25, trench_width=6, mask_th=2,
                            sub_top=18.0, hole=True)
    m = petch.SF6O2(meanE=150, iedf="mean")
    assert m.par['Emean'] == 150 and m.par['ied_mode'] == "mean" and m.par['rate_scale'] == 0.025
    assert m.flags['ion_reflection'] and m.flags['chemistry'] == "belen"


def test_faces_in_mask_host_predicate():
    from petch.threed import faces_in_mask
    geo = dict(sub_top=18.0, Lx=14.0, Ly=14.0)
    c = np.array([[7, 7, 19], [0, 0, 19], [0, 0, 20], [0, 0, 18], [10, 7, 19.5]], float)
    np.testing.assert_array_equal(faces_in_mask(c, geo, 2.0, 6.0, hole=True),
                                  oc.faces_in_mask(c, geo, 2.0, 6.0, hole=True))
    np.testing.assert_array_equal(faces_in_mask(c, geo, 2.0, 6.0, hole=True),
                                  [False, True, False, True, True])   # r == w/2 is not open


def test_gpu_path_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from petch import threed
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        threed.run_etch_3d(Lx=3.0, Ly=0.9, Lz=6.5, dx=0.15, trench_width=0.9, mask_th=0.3,
                           sub_top=6.0, n_steps=1, verbose=False)
    from petch._abi import Context
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        Context(0)

