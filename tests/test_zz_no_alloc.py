"""petch_reinit_fsm_noalloc (the reinit with its workspace in a fixed per-GPU device array, for
callers that cannot allocate) against petch_reinit_fsm through the C-ABI: bit-identical output
on two calls in a row, the second on a smaller grid (the workspace keeps the first call's
contents), and a clear capacity error for a grid beyond the fixed workspace."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _phi(nx, ny, nz, dx, seed):
    rng = np.random.default_rng(seed)
    x, y, z = (np.arange(n) * dx for n in (nx, ny, nz))
    X, Y, Z = np.meshgrid(x, y, z, indexing='ij')
    c = np.array([nx, ny, nz]) * dx / 2
    r = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
    # a sphere's distance, distorted so the reinit has work to do
    phi = (r - 0.3 * min(nx, ny, nz) * dx) * (1.0 + 0.4 * np.sin(X) * np.cos(Y))
    return (phi + 0.01 * dx * rng.normal(size=phi.shape)).astype(np.float32)


def test_noalloc_reinit_equals_reinit_on_two_calls():
    import torch
    from petch._abi import Context, ptr
    dev = torch.device("cuda", 0)
    ctx = Context.get(dev)
    dx, band = 0.1, 0.6
    for k, (nx, ny, nz) in enumerate([(96, 80, 120), (40, 36, 52)]):
        print(f"call {k}: grid {nx}x{ny}x{nz}", flush=True)
        phi = torch.from_numpy(_phi(nx, ny, nz, dx, k)).to(dev)
        a, b = phi.clone(), phi.clone()
        ctx.call("petch_reinit_fsm", ptr(a), nx, ny, nz, dx, band, -1)
        ctx.call("petch_reinit_fsm_noalloc", ptr(b), nx, ny, nz, dx, band, -1)
        torch.cuda.synchronize()
        assert torch.equal(a, b), float((a - b).abs().max())
        assert float((a - phi).abs().max()) > 0.0                   # it did reinitialise


def test_noalloc_reinit_rejects_a_grid_beyond_its_workspace():
    import torch
    from petch._abi import Context, PetchCapacityError, ptr
    dev = torch.device("cuda", 0)
    ctx = Context.get(dev)
    phi = torch.zeros(8, device=dev)
    # the size check runs before any launch, so the short buffer is never touched
    with pytest.raises(PetchCapacityError, match="fixed workspace"):
        ctx.call("petch_reinit_fsm_noalloc", ptr(phi), 512, 512, 512, 0.1, 0.6, -1)
