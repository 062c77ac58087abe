"""ctypes binding of libpetch_hip.so (declared in include/petch_hip.h).

Device memory is torch tensors on the ROCm device; every call passes raw device pointers and
enqueues on torch's current stream. There is no CPU fallback: if the library or a GPU is
missing, the first call raises.
"""
import ctypes
import os
from ctypes import c_int, c_double, c_float, c_void_p, c_char_p, c_uint64, POINTER

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PETCH_HIP_LIB", os.path.join(_HERE, "libpetch_hip.so"))

PETCH_OK = 0
PETCH_E_ARG = -1
PETCH_E_HIP = -2
PETCH_E_CAPACITY = -3
PETCH_E_NONFINITE = -4
PETCH_E_STATE = -5


class PetchError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class PetchArgError(PetchError, ValueError):
    pass


class PetchCapacityError(PetchError):
    pass


class PetchNonFiniteError(PetchError):
    pass


_ERR = {PETCH_E_ARG: PetchArgError, PETCH_E_CAPACITY: PetchCapacityError,
        PETCH_E_NONFINITE: PetchNonFiniteError}

P = c_void_p


class Box(ctypes.Structure):
    """petch_box: ray-tracer box (source plane z_src, lateral periodic wrap flag)."""
    _fields_ = [("Lx", c_float), ("Ly", c_float), ("Lz", c_float), ("z_src", c_float),
                ("periodic", c_int)]


class Source(ctypes.Structure):
    """petch_source: 0 neutral cosine, 1 ion Gaussian tilt, 2 ion power-cosine."""
    _fields_ = [("kind", c_int), ("sigma", c_float), ("exponent", c_float)]


class ChemParams(ctypes.Structure):
    """petch_chem_params."""
    _fields_ = [(k, c_double) for k in ("ionFlux", "Fflux", "Oflux", "k_sigma", "beta_sigma",
                                        "rho", "rate_scale", "B_sp", "fnorm_E", "fnorm_O", "Yie",
                                        "Ysp", "Yp", "s_F", "s_O")] + \
              [("angular_viennaps", c_int), ("chemistry_belen", c_int)]


class MaskGeom(ctypes.Structure):
    """petch_mask_geom."""
    _fields_ = [(k, c_double) for k in ("sub_top", "mask_th", "Lx", "Ly", "trench_width")] + \
              [("hole", c_int)]


class IonParams(ctypes.Structure):
    _fields_ = [(k, c_float) for k in ("meanE", "sigmaE", "A_sp", "Eth_sp", "B_sp", "A_ie",
                                       "Eth_ie", "A_p", "Eth_p", "inflect", "minang", "n_l",
                                       "A_en", "minE", "thrmin", "thrmax")]

class NeutralFeature(ctypes.Structure):
    """petch_neutral_feature: one feature of a batched Russian-roulette neutral launch."""
    _fields_ = [("bvh", c_void_p), ("box", Box), ("bare", c_void_p),
                ("beta_a", c_float), ("n_a", c_int), ("seed_a", c_uint64), ("fl_a", c_void_p),
                ("beta_b", c_float), ("n_b", c_int), ("seed_b", c_uint64), ("fl_b", c_void_p)]


class IonFeature(ctypes.Structure):
    """petch_ion_feature: one feature of a batched faithful-ion launch."""
    _fields_ = [("bvh", c_void_p), ("box", Box), ("src", Source), ("ip", IonParams),
                ("n", c_int), ("seed", c_uint64), ("sp", c_void_p), ("ie", c_void_p),
                ("ipass", c_void_p), ("arr", c_void_p), ("ang", c_void_p)]


class AdvectItem(ctypes.Structure):
    """petch_advect_item: one grid of a batched advection."""
    _fields_ = [("phi", c_void_p), ("work", c_void_p), ("F", c_void_p), ("mask", c_void_p),
                ("mask_phi", c_void_p), ("nx", c_int), ("ny", c_int), ("nz", c_int),
                ("dx", c_double), ("dt", c_double), ("nsub", c_int)]


class ReinitItem(ctypes.Structure):
    """petch_reinit_item: one grid of a batched reinit."""
    _fields_ = [("phi", c_void_p), ("nx", c_int), ("ny", c_int), ("nz", c_int), ("dx", c_double),
                ("band", c_double), ("n_iter", c_int)]


class RoundItem(ctypes.Structure):
    """petch_round_item: one feature of a batched neutral-round finish."""
    _fields_ = [("flF", c_void_p), ("flO", c_void_p), ("area", c_void_p), ("normals", c_void_p),
                ("nbr", c_void_p), ("m_i", c_void_p), ("cos_i", c_void_p), ("iy_ie", c_void_p),
                ("iy_ip", c_void_p), ("m_F", c_void_p), ("m_O", c_void_p), ("bare", c_void_p),
                ("chem", ChemParams), ("nf", c_int), ("smooth", c_int), ("scale", c_float),
                ("hi", c_float), ("alpha", c_float)]


class IonFinishItem(ctypes.Structure):
    """petch_ion_finish_item: one feature of a batched faithful-ion finish."""
    _fields_ = [("acc", c_void_p), ("area", c_void_p), ("normals", c_void_p), ("nbr", c_void_p),
                ("out", c_void_p), ("nf", c_int), ("smooth", c_int), ("scale", c_float),
                ("alpha", c_float)]


class SurfaceItem(ctypes.Structure):
    """petch_surface_item: one mesh of a batched attribute + adjacency build."""
    _fields_ = [("verts", c_void_p), ("faces", c_void_p), ("nf", c_int), ("nv", c_int),
                ("centroids", c_void_p), ("areas", c_void_p), ("normals", c_void_p),
                ("nbr", c_void_p), ("finite", c_void_p)]


class RateItem(ctypes.Structure):
    """petch_rate_item: one surface of a batched surface-rate launch."""
    _fields_ = [(k, c_void_p) for k in ("m_i", "m_F", "m_O", "cos_i", "iy_sp", "iy_ie", "iy_ip",
                                        "centroids", "is_mask")] + \
               [("mg", MaskGeom), ("chem", ChemParams), ("F", c_int), ("V", c_void_p),
                ("red", c_void_p), ("bvh", c_void_p), ("bvh_err", c_void_p)]


class KnudsenParams(ctypes.Structure):
    """petch_knudsen_params."""
    _fields_ = [("dx", c_double), ("sub_top", c_double), ("Ly", c_double), ("trench", c_int),
                ("wall_loss_scale", c_double), ("front_loss", c_int), ("band_W", c_double),
                ("pass_frac", c_double), ("ar_pass", c_double), ("footprint", c_int),
                ("cx", c_double), ("cy", c_double), ("half_width", c_double),
                ("sink_field", c_int), ("beta_a", c_double), ("beta_b", c_double)]


class DdaParams(ctypes.Structure):
    """petch_dda_params."""
    _fields_ = [("dx", c_double), ("n_dir", c_int), ("n_reemit", c_int), ("periodic_y", c_int),
                ("beta_a", c_double), ("beta_b", c_double)]


class McItem(ctypes.Structure):
    """petch_mc_item: one grid of a batched marching cubes."""
    _fields_ = [("phi", c_void_p), ("nx", c_int), ("ny", c_int), ("nz", c_int), ("dx", c_double),
                ("verts", c_void_p), ("faces", c_void_p), ("cap_v", c_int), ("cap_f", c_int),
                ("nv_out", c_void_p), ("nf_out", c_void_p)]


class ExtendItem(ctypes.Structure):
    """petch_extend_item: one grid of a batched velocity extension."""
    _fields_ = [("bvh", c_void_p), ("V", c_void_p), ("phi", c_void_p), ("Fs", c_void_p),
                ("nx", c_int), ("ny", c_int), ("nz", c_int), ("dx", c_double), ("band", c_double)]


class NearestItem(ctypes.Structure):
    """petch_nearest_item: one point set of a batched nearest-face lookup."""
    _fields_ = [("bvh", c_void_p), ("src", c_void_p), ("pts", c_void_p), ("n", c_int),
                ("miss", c_float), ("out", c_void_p)]


# name -> argument types (all entry points return int)
SIGNATURES = {
    "petch_ctx_create": [c_int, POINTER(c_void_p)],
    "petch_ctx_destroy": [P],
    "petch_set_stream": [P, P],
    "petch_synchronize": [P],
    "petch_set_deterministic": [P, c_int],
    "petch_version": [],
    "petch_init_sdf": [P, P, P, c_int, c_int, c_int, c_double, c_double, c_double, c_double,
                       c_double, c_double, c_int],
    "petch_advect": [P, P, P, P, P, P, c_int, c_int, c_int, c_double, c_double, c_int],
    "petch_reinit_fsm": [P, P, c_int, c_int, c_int, c_double, c_double, c_int],
    "petch_reinit_fsm_noalloc": [P, P, c_int, c_int, c_int, c_double, c_double, c_int],
    "petch_profile_area": [P, P, c_int, c_int, c_int, c_double, c_double, P],
    "petch_advect_batch": [P, c_int, P],
    "petch_bvh_build_batch": [P, c_int, P, P, P, P, P],
    "petch_reinit_fsm_batch": [P, c_int, P],
    "petch_reinit_fsm_adjoint": [P, P, c_int, c_int, c_int, c_double, c_double, c_int, P, P],
    "petch_profile_area_adjoint": [P, P, c_int, c_int, c_int, c_double, c_double, P, P],
    "petch_marching_cubes": [P, P, c_int, c_int, c_int, c_double, P, P, c_int, c_int,
                             POINTER(c_int), POINTER(c_int)],
    "petch_mesh_attrs": [P, P, P, c_int, P, P, P],
    "petch_face_adjacency": [P, P, c_int, c_int, P],
    "petch_gas_normals": [P, P, c_int, c_int, c_int, c_double, P, P, c_int, P],
    "petch_bvh_create": [P, POINTER(c_void_p)],
    "petch_bvh_destroy": [P, P],
    "petch_bvh_build": [P, P, P, P, c_int, c_int],
    "petch_bvh_set_quality": [P, P, c_int],
    "petch_set_trace_occupancy": [P, c_int],
    "petch_trace_occupancy_limit": [P, POINTER(c_int)],
    "petch_bvh_query_ray": [P, P, P, P, c_int, c_float, P, P],
    "petch_bvh_query_point": [P, P, P, c_int, c_float, P, P],
    "petch_bvh_status": [P, P, POINTER(c_int)],
    "petch_bvh_status_async": [P, P, P],
    "petch_bvh_root_bounds": [P, P, P],
    "petch_gen_source": [P, POINTER(Box), POINTER(Source), c_int, c_uint64, P, P],
    "petch_trace_neutral_cap": [P, P, POINTER(Box), POINTER(Source), P, c_float, c_int, c_int,
                                c_uint64, P],
    "petch_trace_redeposition": [P, P, P, P, P, P, c_float, c_int, c_uint64, P],
    "petch_trace_neutral_rr": [P, P, POINTER(Box), POINTER(Source), P, c_float, c_int, c_uint64,
                               P],
    "petch_trace_neutral_rr_adjoint": [P, P, POINTER(Box), POINTER(Source), P, c_float, c_int,
                                       c_uint64, P, P],
    "petch_trace_neutral_rr_pair_adjoint": [P, P, POINTER(Box), POINTER(Source), P, c_float,
                                            c_int, c_uint64, P, c_float, c_int, c_uint64, P, P],
    "petch_trace_neutral_rr_pair": [P, P, POINTER(Box), POINTER(Source), P, c_float, c_int,
                                    c_uint64, P, c_float, c_int, c_uint64, P],
    "petch_trace_neutral_rr_batch": [P, POINTER(Source), c_int, P],
    "petch_trace_ion_yield_batch": [P, c_int, P],
    "petch_trace_ion_yield": [P, P, POINTER(Box), POINTER(Source), POINTER(IonParams), c_int,
                              c_uint64, P, P, P, P, P],
    "petch_trace_ion_yield_tangent": [P, P, POINTER(Box), POINTER(Source), POINTER(IonParams),
                                      c_int, c_uint64, P, P, P],
    "petch_trace_legacy": [P, P, POINTER(Box), POINTER(Source), c_float, c_int, c_int, c_float,
                           c_float, c_int, c_uint64, P, P],
    "petch_trace_ff": [P, P, POINTER(Box), P, P, c_int, c_int, c_uint64, P],
    "petch_trace_hit_counts": [P, c_int, P],
    "petch_normalize": [P, P, P, c_int, c_float, c_float, c_float, P],
    "petch_cos_ratio": [P, P, P, c_int, P],
    "petch_smooth": [P, P, P, P, c_int, c_int, c_float, c_int, P],
    "petch_normalize_adjoint": [P, P, P, c_int, c_float, c_float, c_float, P, P],
    "petch_smooth_adjoint": [P, P, P, c_int, c_int, c_float, c_int, P, P],
    "petch_belen_coverage": [P, P, P, P, P, P, P, POINTER(ChemParams), c_int, P, P, P],
    "petch_finish_round_batch": [P, c_int, P],
    "petch_finish_ion_batch": [P, c_int, P],
    "petch_surface_build_batch": [P, c_int, P],
    "petch_surface_rate_batch": [P, c_int, P],
    "petch_gaussian_smooth": [P, P, c_int, c_int, c_int, P, c_int, P, P],
    "petch_reinit_pde": [P, P, c_int, c_int, c_int, c_double, c_int],
    "petch_knudsen_face_flux": [P, P, c_int, c_int, c_int, P, P, P, P, c_int,
                                POINTER(KnudsenParams), P, P],
    "petch_dda_neutral_flux": [P, P, c_int, c_int, c_int, P, P, P, P, c_int,
                               POINTER(DdaParams), P, P],
    "petch_belen_coverage_adjoint": [P, P, P, P, P, P, P, POINTER(ChemParams), c_int, P, P, P,
                                     P, P, P],
    "petch_surface_rate": [P, P, P, P, P, P, P, P, P, P, POINTER(MaskGeom),
                           POINTER(ChemParams), c_int, P, P],
    "petch_extend": [P, P, P, P, c_int, c_int, c_int, c_double, c_double, P],
    "petch_extend_adjoint": [P, P, P, c_int, c_int, c_int, c_double, c_double, P, P],
    "petch_surface_rate_adjoint": [P, P, P, P, P, P, P, P, P, P, POINTER(MaskGeom),
                                   POINTER(ChemParams), c_int, P, P, P, P, P, P, P, P],
    "petch_nearest_face_value": [P, P, P, P, c_int, c_float, P],
    "petch_extend_batch": [P, c_int, P],
    "petch_advect_step": [P, P, P, P, P, c_int, c_int, c_int, c_double, c_double, P],
    "petch_advect_adjoint": [P, P, P, P, P, c_int, c_int, c_int, c_double, c_double, P, P],
    "petch_marching_cubes_batch": [P, c_int, P],
    "petch_nearest_face_value_batch": [P, c_int, P],
    "petch_radiosity_assemble": [P, P, c_int, c_int, P, P, P],
    "petch_radiosity_solve": [P, P, P, P, c_int, c_int, P, c_double, c_int, c_double, P, P,
                              POINTER(c_int)],
    "petch_radiosity_adjoint": [P, P, c_int, c_int, P, P, c_double, P, c_int, c_double, P, P,
                                POINTER(c_int)],
}

_lib = None


def lib():
    """Loads libpetch_hip.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libpetch_hip.so not found at {LIB_PATH}; build it with "
                "`python plasma-etching-code_amd/build.py` (the HIP path has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = c_int
        L.petch_last_error.argtypes = [P]
        L.petch_last_error.restype = c_char_p
        _lib = L
    return _lib


class Context:
    """One petch_ctx per device; kernels go to torch's current stream on that device."""
    _by_device = {}

    def __init__(self, device):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("petch HIP path: no ROCm GPU visible (torch.cuda.is_available() "
                               "is False); the HIP implementation has no CPU fallback")
        self.device = int(device)
        h = c_void_p()
        rc = lib().petch_ctx_create(self.device, ctypes.byref(h))
        if rc != PETCH_OK:
            raise PetchError(rc, f"petch_ctx_create(device={self.device}) failed")
        self.h = h

    @classmethod
    def get(cls, device=None):
        import torch
        if device is None:
            device = torch.cuda.current_device()
        if isinstance(device, torch.device):
            device = device.index if device.index is not None else torch.cuda.current_device()
        c = cls._by_device.get(int(device))
        if c is None:
            c = cls(device)
            cls._by_device[int(device)] = c
        return c

    def call(self, name, *args):
        import torch
        L = lib()
        L.petch_set_stream(self.h, c_void_p(torch.cuda.current_stream(self.device).cuda_stream))
        rc = getattr(L, name)(self.h, *args)
        if rc != PETCH_OK:
            msg = L.petch_last_error(self.h)
            msg = msg.decode() if msg else ""
            raise _ERR.get(rc, PetchError)(rc, f"{name}: {msg}")
        return rc


def ptr(t, dtype=None, name="tensor"):
    """Raw device pointer of a contiguous ROCm tensor (None -> NULL)."""
    if t is None:
        return c_void_p(0)
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    return c_void_p(t.data_ptr())
