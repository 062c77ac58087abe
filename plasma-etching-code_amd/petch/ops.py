"""Tensor-level wrappers over the C ABI: each takes/returns torch tensors on the ROCm device and
enqueues on torch's current stream. The reference-shaped numpy API in threed.py is built on
these; the etch driver calls them directly so state never leaves the GPU between stages.
"""
import torch

from ._abi import Context, ptr

f32 = torch.float32
u8 = torch.uint8
i32 = torch.int32


def _ctx(t):
    return Context.get(t.device)


def _shape3(phi):
    if phi.dim() != 3:
        raise ValueError(f"level set must be 3-D (nx,ny,nz), got shape {tuple(phi.shape)}")
    return [int(s) for s in phi.shape]


def init_sdf(nx, ny, nz, dx, Lx, Ly, trench_width, mask_th, sub_top, hole, device=None):
    """Initial phi (float32) and mask (uint8) of make_trench_3d (threed.py:32-59)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    phi = torch.empty((nx, ny, nz), dtype=f32, device=dev)
    mask = torch.empty((nx, ny, nz), dtype=u8, device=dev)
    _ctx(phi).call("petch_init_sdf", ptr(phi), ptr(mask), nx, ny, nz, float(dx), float(Lx),
                   float(Ly), float(trench_width), float(mask_th), float(sub_top), int(bool(hole)))
    return phi, mask


def advect(phi, F, mask, mask_phi, dx, dt, nsub, work=None):
    """nsub upwind substeps of size dt with mask re-pin, in place on phi (threed.py:1295-1307)."""
    nx, ny, nz = _shape3(phi)
    for t, n, d in ((F, "F", f32), (mask_phi, "mask_phi", f32), (mask, "mask", u8)):
        if tuple(t.shape) != (nx, ny, nz):
            raise ValueError(f"advect: {n} shape {tuple(t.shape)} != phi shape {(nx, ny, nz)}")
    if work is None:
        work = torch.empty_like(phi)
    _ctx(phi).call("petch_advect", ptr(phi, f32, "phi"), ptr(work, f32, "work"), ptr(F, f32, "F"),
                   ptr(mask, u8, "mask"), ptr(mask_phi, f32, "mask_phi"), nx, ny, nz, float(dx),
                   float(dt), int(nsub))
    return phi


def reinit_fsm(phi, dx, band, n_iter=None):
    """Jacobi Godunov-Eikonal reinit in place (threed.py:1515-1537)."""
    nx, ny, nz = _shape3(phi)
    _ctx(phi).call("petch_reinit_fsm", ptr(phi, f32, "phi"), nx, ny, nz, float(dx), float(band),
                   -1 if n_iter is None else int(n_iter))
    return phi
