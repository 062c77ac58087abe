/* petch_hip.h — C ABI of libpetch_hip.so, the MI355X (gfx950) implementation of petch's
 * per-step 3-D etch loop.
 *
 * Conventions shared by every entry point:
 *   - All array arguments are DEVICE pointers (e.g. torch.Tensor.data_ptr() on a ROCm device)
 *     unless the parameter name ends in `_host`.
 *   - Level-set grids are (nx, ny, nz) float32 in C order (z fastest), node (i,j,k) at
 *     (i*dx, j*dx, k*dx); phi > 0 is solid, phi < 0 is gas (reference threed.py:32-59).
 *   - Every kernel is enqueued on the stream set by petch_set_stream (default: the null stream);
 *     entry points do not synchronise unless their comment says they return a count to the host.
 *   - Return value: PETCH_OK (0) or a negative PETCH_E_* code; petch_last_error() gives the text.
 *
 * Each entry point names the reference function it replaces (file:line in
 * sdelaurentiis123/plasma-etching-code, src/petch/...).
 */
#ifndef PETCH_HIP_H
#define PETCH_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PETCH_OK 0
#define PETCH_E_ARG (-1)       /* invalid argument (shape, size, null pointer) */
#define PETCH_E_HIP (-2)       /* HIP runtime error (launch, allocation, copy) */
#define PETCH_E_CAPACITY (-3)  /* output capacity too small; the required size is reported */
#define PETCH_E_NONFINITE (-4) /* non-finite input where a finite one is required */
#define PETCH_E_STATE (-5)     /* object used before it was built */

typedef struct petch_ctx petch_ctx;
typedef struct petch_bvh petch_bvh;

/* ---------------------------------------------------------------- context ---------------- */
/* Replaces the implicit Warp runtime state (threed.py:22-23 wp.init / DEVICE). */
int petch_ctx_create(int device, petch_ctx** out);
int petch_ctx_destroy(petch_ctx* ctx);
/* `stream` is a hipStream_t (torch.cuda.current_stream().cuda_stream); 0 = null stream. */
int petch_set_stream(petch_ctx* ctx, void* stream);
int petch_synchronize(petch_ctx* ctx);
const char* petch_last_error(petch_ctx* ctx);
int petch_version(void);

/* ---------------------------------------------------------------- a1: initial SDF -------- */
/* make_trench_3d (threed.py:32-59): phi = exact signed distance to the CSG solid
 * {z < sub_top} U ({sub_top <= z < sub_top+mask_th} minus the opening), mask = mask material
 * nodes (threed.py:37-43, evaluated in float64 exactly as numpy does). Writes phi (float32) and
 * mask (uint8, 1 = mask). hole != 0: circular opening; else a y-invariant trench. */
int petch_init_sdf(petch_ctx* ctx, float* phi, uint8_t* mask, int nx, int ny, int nz,
                   double dx, double Lx, double Ly, double trench_width, double mask_th,
                   double sub_top, int hole);

/* ---------------------------------------------------------------- a19: advection --------- */
/* advect_3d_gpu + _advect_iter (threed.py:1262-1307): nsub first-order upwind Godunov substeps
 * phi <- phi - dt*F*|grad phi|, one-sided at the domain boundary, mask nodes re-pinned to
 * mask_phi after every substep. phi is updated in place; `work` is an (nx,ny,nz) float32
 * buffer used for ping-pong. dt is the substep size (dt_step/nsub). */
int petch_advect(petch_ctx* ctx, float* phi, float* work, const float* F, const uint8_t* mask,
                 const float* mask_phi, int nx, int ny, int nz, double dx, double dt, int nsub);

/* ---------------------------------------------------------------- a20: FSM reinit -------- */
/* reinit_fsm (threed.py:1515-1537) with _eik_init / _eik_jacobi / _eik_solve
 * (threed.py:1421-1512): big = band+4dx, n_iter = ceil(band/dx)+4 when n_iter < 0, output
 * sign(phi0)*min(U, band+dx) with sign(0) = +1. phi is replaced in place. */
int petch_reinit_fsm(petch_ctx* ctx, float* phi, int nx, int ny, int nz, double dx,
                     double band, int n_iter);

#ifdef __cplusplus
}
#endif
#endif /* PETCH_HIP_H */
