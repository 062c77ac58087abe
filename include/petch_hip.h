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

/* Simulation box of the ray tracer: source plane z = z_src (Lz - dx in the reference,
 * threed.py:477,749), lateral periodic wrap when periodic != 0 (par['periodic_y'],
 * threed.py:289-349). */
typedef struct petch_box {
  float Lx, Ly, Lz, z_src;
  int periodic;
} petch_box;

/* Source distribution (threed.py:410-458): kind 0 = neutral, cosine hemisphere about -z;
 * 1 = ion, Gaussian tilt of std `sigma` rad per axis; 2 = ion, power-cosine cos^exponent(theta)
 * about -z. */
typedef struct petch_source {
  int kind;
  float sigma;
  float exponent;
} petch_source;

/* Faithful-ion constants of _trace3d_ion_yield (threed.py:600-608, 800-802, 806-813). */
typedef struct petch_ion_params {
  float meanE, sigmaE, A_sp, Eth_sp, B_sp, A_ie, Eth_ie, A_p, Eth_p;
  float inflect, minang, n_l, A_en, minE, thrmin, thrmax;
} petch_ion_params;

/* Chemistry constants (PAR, params.py:13-76) with the energy-integrated yields Yie/Ysp/Yp of
 * chemistry._yields (chemistry.py:34-39, evaluated on the host); angular_viennaps selects
 * Flags.yield_angular == 'viennaps', chemistry_belen selects Flags.chemistry == 'belen'. */
typedef struct petch_chem_params {
  double ionFlux, Fflux, Oflux, k_sigma, beta_sigma, rho, rate_scale, B_sp, fnorm_E, fnorm_O;
  double Yie, Ysp, Yp, s_F, s_O;
  int angular_viennaps, chemistry_belen;
} petch_chem_params;

/* Mask geometry of faces_in_mask (threed.py:1608-1616). */
typedef struct petch_mask_geom {
  double sub_top, mask_th, Lx, Ly, trench_width;
  int hole;
} petch_mask_geom;

/* ---------------------------------------------------------------- context ---------------- */
/* Replaces the implicit Warp runtime state (threed.py:22-23 wp.init / DEVICE). */
int petch_ctx_create(int device, petch_ctx** out);
int petch_ctx_destroy(petch_ctx* ctx);
/* `stream` is a hipStream_t (torch.cuda.current_stream().cuda_stream); 0 = null stream. */
int petch_set_stream(petch_ctx* ctx, void* stream);
int petch_synchronize(petch_ctx* ctx);
/* Deterministic mode of the context's ray traces (on = 1): the per-face accumulators of
 * petch_trace_neutral_rr[_pair|_batch] and petch_trace_ion_yield[_batch] are summed as 64-bit
 * integers in units of 2^-32 instead of float atomics, then converted to float, so a trace's
 * output no longer depends on the order in which the GPU schedules its rays: the same inputs
 * give bit-identical outputs, and with them a whole etch run repeats exactly. The sums are the
 * exact sums of the deposits rounded to 2^-32 (the float path rounds every add to float). No
 * reference counterpart (the reference accumulates with float atomics, threed.py:579,650-654).
 * Default off. */
int petch_set_deterministic(petch_ctx* ctx, int on);
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
/* One advection substep of petch_advect out of place: out = the substep from phi (out != phi).
 * No reference counterpart: the differentiable advection keeps every substep's state with it. */
int petch_advect_step(petch_ctx* ctx, const float* phi, const float* F, const uint8_t* mask,
                      const float* mask_phi, int nx, int ny, int nz, double dx, double dt,
                      float* out);
/* Adjoint of one advection substep (petch_advect with nsub = 1) at the state phi it started
 * from: gphi = dL/dphi for gout = dL/d(output), and gF += dL/dF. The upwind choices are those
 * of phi; where |grad phi|_upwind = 0 the sqrt term contributes no derivative; mask nodes pass
 * no gradient. gphi must not alias phi, gout or gF. Two launches (each node's own stencil,
 * then what its neighbours' stencils owe it) through a grid-sized scratch buffer. No reference
 * counterpart: the reference's etch loop is not differentiated (its only gradients are demo
 * tapes, scripts/inverse_design.py:95-101). */
int petch_advect_adjoint(petch_ctx* ctx, const float* phi, const float* F, const uint8_t* mask,
                         const float* gout, int nx, int ny, int nz, double dx, double dt,
                         float* gphi, float* gF);

/* The substeps of petch_advect for many grids at once (the features of a batched step, no
 * reference counterpart: the reference advances one feature per run_etch_3d): one launch per
 * substep index for all grids (a grid with fewer substeps sits out), then one copy-back launch
 * for the grids whose last substep ended in `work`. Each item's result equals petch_advect's. */
typedef struct petch_advect_item {
  float* phi;
  float* work;               /* a grid-sized buffer, distinct from phi */
  const float* F;
  const uint8_t* mask;
  const float* mask_phi;
  int nx, ny, nz;
  double dx, dt;             /* dt is the substep's step */
  int nsub;
} petch_advect_item;
int petch_advect_batch(petch_ctx* ctx, int n, const petch_advect_item* items_host);

/* ---------------------------------------------------------------- a20: FSM reinit -------- */
/* reinit_gpu (threed.py:1399-1418), reinit_method='gpu': n_iter forward-Euler steps
 * (dtau = 0.3 dx) of phi_tau + sgn(phi0)(|grad phi| - 1) = 0 in float32 with the Russo-Smereka
 * interface fix (the kernel _reinit_iter, threed.py:1349-1398), in place; the reference runs 24. */
int petch_reinit_pde(petch_ctx* ctx, float* phi, int nx, int ny, int nz, double dx, int n_iter);

/* gaussian_filter(phi, sigma) of run_etch_3d's surf_smooth (threed.py:1779-1782, scipy.ndimage,
 * mode 'reflect'), in place: three axis passes (x, y, z) of the 2 radius + 1 weights (host
 * array, the normalised exp(-x^2 / 2 sigma^2) as scipy forms them), each in float64 stored as
 * float32; then nodes with mask[n] != 0 take mask_phi[n] (mask may be NULL). */
int petch_gaussian_smooth(petch_ctx* ctx, float* phi, int nx, int ny, int nz,
                          const double* weights_host, int radius, const uint8_t* mask,
                          const float* mask_phi);

/* reinit_fsm (threed.py:1515-1537) with _eik_init / _eik_jacobi / _eik_solve
 * (threed.py:1421-1512): big = band+4dx, n_iter = ceil(band/dx)+4 when n_iter < 0, output
 * sign(phi0)*min(U, band+dx) with sign(0) = +1. phi is replaced in place. */
int petch_reinit_fsm(petch_ctx* ctx, float* phi, int nx, int ny, int nz, double dx,
                     double band, int n_iter);

/* petch_reinit_fsm for many grids at once (the features of a batched step; no reference
 * counterpart): one launch per stage (init, each Jacobi sweep, final) for all grids, each grid
 * with its own dx, band and sweep count (n_iter < 0: ceil(band/dx)+4). Each item's result
 * equals petch_reinit_fsm's. */
typedef struct petch_reinit_item {
  float* phi;
  int nx, ny, nz;
  double dx, band;
  int n_iter;
} petch_reinit_item;
int petch_reinit_fsm_batch(petch_ctx* ctx, int n, const petch_reinit_item* items_host);

/* Adjoint of petch_reinit_fsm (no reference counterpart: the reference's etch loop is not
 * differentiated): gphi = d loss / d phi0 for gout = d loss / d (reinit output), phi0 the input
 * level set. The exact adjoint of the piecewise-smooth map the reinit computes, every discrete
 * choice (interface flags, the smaller neighbour per axis, the active Godunov terms, whether a
 * sweep lowers a node, the clamp at band + dx) held at the forward's; the sweep states are
 * rebuilt with the forward's kernels (workspace (n_iter + 1) * 4 + 25 bytes per node). */
int petch_reinit_fsm_adjoint(petch_ctx* ctx, const float* phi0, int nx, int ny, int nz,
                             double dx, double band, int n_iter, const float* gout, float* gphi);

/* petch_reinit_fsm without any allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops that forbid hipMalloc). No counterpart in the
 * reference. The same kernels and results as petch_reinit_fsm; the workspace (U, U2 and the
 * frozen flags, about 9 bytes per node) is a fixed 64 MiB device array per GPU, resolved when
 * the context is created, so grids are limited to about 7.4M nodes (PETCH_E_CAPACITY beyond).
 * Every call on one GPU shares that GPU's workspace: calls that overlap on one GPU (two
 * streams, two host threads) race and must be serialised by the caller. */
int petch_reinit_fsm_noalloc(petch_ctx* ctx, float* phi, int nx, int ny, int nz, double dx,
                             double band, int n_iter);


/* ---------------------------------------------------------------- a2: marching cubes ----- */
/* extract_mesh_3d / extract_mesh_3d_gpu (threed.py:62-100): iso-0 surface of phi, vertices in
 * physical units (float32, (nv,3)), faces int32 (nf,3), normals of cross(v1-v0,v2-v0) pointing
 * into the gas. Deterministic order: vertices by (node, axis) of their grid edge, faces by cell.
 * Synchronises once to return nv / nf to the host. If nv > cap_v or nf > cap_f nothing is
 * written and PETCH_E_CAPACITY is returned with *nv_host / *nf_host set to the sizes needed
 * (pass NULL buffers and zero capacities to query the sizes). */
int petch_marching_cubes(petch_ctx* ctx, const float* phi, int nx, int ny, int nz, double dx,
                         float* verts, int* faces, int cap_v, int cap_f, int* nv_host,
                         int* nf_host);

/* petch_marching_cubes for many grids with one host read (the batched etch loop): every grid's
 * counts come back in one copy, then every grid's surface is emitted. nv_out / nf_out are host
 * pointers; on PETCH_E_CAPACITY every item's counts are set and nothing was written. No
 * reference counterpart: the reference meshes one feature per call. */
typedef struct petch_mc_item {
  const float* phi;
  int nx, ny, nz;
  double dx;
  float* verts;
  int* faces;
  int cap_v, cap_f;
  int* nv_out;
  int* nf_out;
} petch_mc_item;
int petch_marching_cubes_batch(petch_ctx* ctx, int n, const petch_mc_item* items);

/* Face centroids (nf,3) = vertex mean, areas (nf) = 0.5|cross|, unit normals (nf,3) (zero for a
 * degenerate face) — threed.py:96-99 and _face_normals_k threed.py:228-234. Any output may be
 * NULL. */
int petch_mesh_attrs(petch_ctx* ctx, const float* verts, const int* faces, int nf,
                     float* centroids, float* areas, float* normals);

/* _edge_adjacency (threed.py:103-125) as a per-face table: nbr[3f+e] = the other face sharing
 * edge e of face f (e: v0v1, v1v2, v2v0), or -1 on an open boundary. */
int petch_face_adjacency(petch_ctx* ctx, const int* faces, int nf, int nv, int* nbr);

/* petch_mesh_attrs and petch_face_adjacency for many meshes at once (the features of a batched
 * step; no reference counterpart): one launch per stage for all meshes, the outputs those two
 * calls give (centroids, areas, unit normals, and the 3-per-face edge neighbours). finite (may
 * be NULL) receives 1 if every centroid is finite, else 0: the warm-start guard of
 * threed.py:1699, read for all features at once. */
typedef struct petch_surface_item {
  const float* verts;
  const int* faces;
  int nf, nv;
  float *centroids, *areas, *normals;
  int* nbr;
  int* finite;
} petch_surface_item;
int petch_surface_build_batch(petch_ctx* ctx, int n, const petch_surface_item* items_host);

/* _gas_normals (threed.py:1553-1568): unit normals flipped to agree with -grad(phi) (numpy
 * gradient) at the grid node nearest each centroid. */
int petch_gas_normals(petch_ctx* ctx, const float* phi, int nx, int ny, int nz, double dx,
                      const float* centroids, const float* normals, int nf, float* out);

/* ---------------------------------------------------------------- a3-a5: LBVH + queries -- */
/* wp.Mesh(points, indices) (threed.py:1674-1675): a Morton-code LBVH over the triangles,
 * rebuilt from scratch by petch_bvh_build (device buffers grow, never shrink). The handle keeps
 * the mesh's triangles in leaf order, so verts/faces may be overwritten after the build. */
int petch_bvh_create(petch_ctx* ctx, petch_bvh** out);
int petch_bvh_destroy(petch_ctx* ctx, petch_bvh* bvh);
int petch_bvh_build(petch_ctx* ctx, petch_bvh* bvh, const float* verts, const int* faces, int nv,
                    int nf);
/* petch_bvh_build of many meshes (the features of a batched step; no reference counterpart):
 * each tree's parallel stages are launched per tree, the bottom-up refits of all trees run in
 * one launch (a refit is a chain of dependent accesses, tens of microseconds for any tree).
 * Host arrays of n entries; each BVH equals what petch_bvh_build gives. */
int petch_bvh_build_batch(petch_ctx* ctx, int n, petch_bvh* const* bvhs_host,
                          const float* const* verts_host, const int* const* faces_host,
                          const int* nv_host, const int* nf_host);
/* Build quality of the next petch_bvh_build calls: 0 Morton tree collapsed to 4-wide nodes by a
 * fixed rule (fastest build), 1 Morton tree with an area-driven collapse, 2 clustered tree
 * (PLOC) with the area-driven collapse (slowest build, fastest traces), -1 (default) 2 for
 * meshes of 32k faces or more and 0 below. A caller that knows its ray budget picks the
 * trade-off (no reference counterpart: wp.Mesh has one builder). */
int petch_bvh_set_quality(petch_ctx* ctx, petch_bvh* bvh, int quality);
/* wp.mesh_query_ray(mesh, o, d, tmax) (threed.py:306,366,524,569,629,916): closest two-sided hit
 * with 0 < t < tmax (watertight test); face = -1 and t = -1 on a miss. t may be NULL. */
int petch_bvh_query_ray(petch_ctx* ctx, const petch_bvh* bvh, const float* orig,
                        const float* dir, int n, float tmax, int* face, float* t);
/* wp.mesh_query_point_no_sign(mesh, p, max_dist) (threed.py:1328,1626): closest face within
 * max_dist, -1 if none; dist may be NULL. */
int petch_bvh_query_point(petch_ctx* ctx, const petch_bvh* bvh, const float* pts, int n,
                          float max_dist, int* face, float* dist);
/* Traversal error flag (1 = a query overflowed its stack and may have missed a face); syncs. */
int petch_bvh_status(petch_ctx* ctx, const petch_bvh* bvh, int* err_host);
/* The same flag copied into device memory err_dev (int) on the stream, without a host wait, so
 * a caller can gather the flags of many BVHs and read them all at once (no reference
 * counterpart). */
int petch_bvh_status_async(petch_ctx* ctx, const petch_bvh* bvh, int* err_dev);
/* Root AABB (lo xyz, hi xyz) of a BVH with at least two faces; syncs. */
int petch_bvh_root_bounds(petch_ctx* ctx, const petch_bvh* bvh, float* lohi_host);

/* ---------------------------------------------------------------- a6-a10, a21: ray tracing */
/* All traces draw every random number from Philox-4x32-10 keyed by `seed`, one stream per
 * primary ray (counter = (block, ray, 0, 0)); the source ray is the stream's first draws.
 * Per-face outputs are (nf,) float32 and are overwritten. */
/* _gen_source / gen_source_gpu (threed.py:441-469): the source rays the traces use. */
int petch_gen_source(petch_ctx* ctx, const petch_box* box, const petch_source* src, int n,
                     uint64_t seed, float* origin, float* dir);
/* _trace3d_cov_rr (threed.py:549-596): neutral with weight w deposited at every hit,
 * w *= 1 - bare[face]*beta, Russian roulette below 0.1 (survivors w = 0.3), cosine re-emission,
 * <= 1024 bounces. */
int petch_trace_neutral_rr(petch_ctx* ctx, const petch_bvh* bvh, const petch_box* box,
                           const petch_source* src, const float* bare, float beta, int n,
                           uint64_t seed, float* fl);
/* _trace3d_cov (threed.py:511-546), the capped trace of cov_transport = 'cap': every hit
 * deposits 1 into fl; the ray sticks with probability bare[face]*beta (word 0 of its bounce
 * block) and ends at the latest at its (n_reemit + 1)-th hit; cosine re-emission otherwise. No
 * lateral wrap (the reference kernel applies no boundary condition): box->periodic is ignored. */
int petch_trace_neutral_cap(petch_ctx* ctx, const petch_bvh* bvh, const petch_box* box,
                            const petch_source* src, const float* bare, float beta, int n_reemit,
                            int n, uint64_t seed, float* fl);
/* mc_redep_3d (threed.py:1582-1605): n product rays leave face f with probability
 * max(rate_f areas_f, 0) / total (non-finite products count as 0) from centroid + 1e-3 n_f,
 * n_f = gas_normals[f] (unit, into the gas), in the cosine lobe about n_f, then follow the RR
 * walk of petch_trace_neutral_rr with sticking s_redep on every face and no lateral wrap.
 * out (nf,) = deposit * (total / n) / max(areas, 1e-9), the redeposited flux density (all 0 when
 * total <= 0). centroids, gas_normals (nf, 3); areas, rate (nf,). Ray r's face and lobe come
 * from block 0 of its stream, its bounces as in petch_trace_neutral_rr. */
int petch_trace_redeposition(petch_ctx* ctx, const petch_bvh* bvh, const float* centroids,
                             const float* gas_normals, const float* areas, const float* rate,
                             float s_redep, int n, uint64_t seed, float* out);
/* Path-replay adjoint of petch_trace_neutral_rr with the same (bare, beta, n, seed): for
 * gfl = d loss / d fl, gbare (nf, zeroed first) = d loss / d bare with every ray's path and its
 * roulette / termination decisions held fixed (the rays redraw their streams) and a roulette
 * survivor's weight read as w / p with its survival probability p held fixed, so that the
 * expectation of the result is the derivative of the expected flux. A ray totals gfl[f_k] w_k
 * over its path, then adds -beta (total - running total) / (1 - S_i) to gbare[f_i] at each hit,
 * from a buffer of its first 16 hits or by tracing the path again. gbare must not alias bare or
 * gfl. No reference counterpart (the reference does not differentiate _trace3d_cov_rr,
 * threed.py:549-596). */
int petch_trace_neutral_rr_adjoint(petch_ctx* ctx, const petch_bvh* bvh, const petch_box* box,
                                   const petch_source* src, const float* bare, float beta, int n,
                                   uint64_t seed, const float* gfl, float* gbare);
/* petch_trace_neutral_rr_adjoint for the two species of petch_trace_neutral_rr_pair in one
 * launch: gbare (zeroed first) = the sum of both species' gradients (gfl_a for the n_a rays of
 * (beta_a, seed_a), gfl_b for the n_b rays of (beta_b, seed_b)). No reference counterpart. */
int petch_trace_neutral_rr_pair_adjoint(petch_ctx* ctx, const petch_bvh* bvh,
                                        const petch_box* box, const petch_source* src,
                                        const float* bare, float beta_a, int n_a, uint64_t seed_a,
                                        const float* gfl_a, float beta_b, int n_b, uint64_t seed_b,
                                        const float* gfl_b, float* gbare);
/* The two _trace3d_cov_rr launches of one flux/coverage iteration (F then O, threed.py:876-880)
 * as one launch over the same surface and coverage: species a (beta_a, n_a rays, seed_a, fl_a)
 * and species b. Each ray keeps the stream it has in two petch_trace_neutral_rr calls, so the
 * outputs equal theirs up to the order of the float atomic adds; one launch halves the tail of
 * rays still bouncing when the queue runs dry. No reference counterpart of its own. */
int petch_trace_neutral_rr_pair(petch_ctx* ctx, const petch_bvh* bvh, const petch_box* box,
                                const petch_source* src, const float* bare, float beta_a,
                                int n_a, uint64_t seed_a, float* fl_a, float beta_b, int n_b,
                                uint64_t seed_b, float* fl_b);
/* _trace3d_ion_yield (threed.py:599-712): per-ion energy, yield-weighted sputter / ion-enhanced
 * / passivation deposits, arrival weight and weighted cos at every hit, angle-dependent sticking,
 * energy loss and coned-cosine reflection. */
int petch_trace_ion_yield(petch_ctx* ctx, const petch_bvh* bvh, const petch_box* box,
                          const petch_source* src, const petch_ion_params* ip, int n,
                          uint64_t seed, float* sp, float* ie, float* ipass, float* arr,
                          float* ang);
/* Forward-mode derivative of petch_trace_ion_yield (same ip, n, seed, so the same rays and
 * decisions) in the mean ion energy ip->meanE: dsp, die, dip (nf each, zeroed first) =
 * d (sp, ie, ipass) / d meanE per face. Every energy step scales an ion's energy by a factor
 * independent of it, so dE/dmeanE = E / E0 along a path. No reference counterpart. */
int petch_trace_ion_yield_tangent(petch_ctx* ctx, const petch_bvh* bvh, const petch_box* box,
                                  const petch_source* src, const petch_ion_params* ip, int n,
                                  uint64_t seed, float* dsp, float* die, float* dip);
/* _trace3d (threed.py:352-406): specular=0 sticking/diffuse neutral, specular=1 ion with
 * grazing specular reflection (cos_thr, eta); <= n_reemit re-emissions; no periodic wrap. */
int petch_trace_legacy(petch_ctx* ctx, const petch_bvh* bvh, const petch_box* box,
                       const petch_source* src, float sticking, int n_reemit, int specular,
                       float cos_thr, float eta, int n, uint64_t seed, float* flux,
                       float* angacc);
/* ---------------------------------------------------------------- batches of features */
/* Many features (each its own surface, box, coverage and recipe) traced from one ray queue in
 * one launch; feature f's rays carry exactly the streams a single-feature call with the same
 * seeds would give them, so each feature's outputs equal that call's up to the order of the
 * float atomic adds. No reference counterpart: the reference traces one feature per call. */
typedef struct petch_neutral_feature {
  const petch_bvh* bvh;
  petch_box box;
  const float* bare;            /* (nf,) coverage of this feature's faces */
  float beta_a;                 /* species a (F): sticking factor, rays, stream, output */
  int n_a;
  uint64_t seed_a;
  float* fl_a;
  float beta_b;                 /* species b (O) */
  int n_b;
  uint64_t seed_b;
  float* fl_b;
} petch_neutral_feature;

typedef struct petch_ion_feature {
  const petch_bvh* bvh;
  petch_box box;
  petch_source src;
  petch_ion_params ip;
  int n;
  uint64_t seed;
  float* sp;
  float* ie;
  float* ipass;
  float* arr;
  float* ang;
} petch_ion_feature;

/* petch_trace_neutral_rr_pair for every feature of the batch, in one launch (src: the cosine
 * neutral source shared by all features). Outputs are overwritten. */
int petch_trace_neutral_rr_batch(petch_ctx* ctx, const petch_source* src, int nfeat,
                                 const petch_neutral_feature* feats);
/* petch_trace_ion_yield for every feature of the batch, in one launch. */
int petch_trace_ion_yield_batch(petch_ctx* ctx, int nfeat, const petch_ion_feature* feats);

/* Launch configuration of the persistent trace kernels (no reference counterpart: Warp picks
 * its own launch shapes): their grid is CUs x min(blocks_per_cu, the kernel's occupancy limit)
 * workgroups; 0 (the default) means the occupancy limit. The fluxes do not depend on it (a ray's
 * stream is keyed by its id, not by the wavefront that takes it). petch_trace_occupancy_limit
 * reports the limit of the neutral trace kernel. Used by petch.autotune. */
int petch_set_trace_occupancy(petch_ctx* ctx, int blocks_per_cu);
int petch_trace_occupancy_limit(petch_ctx* ctx, int* blocks_per_cu);

/* Ray-hit counters accumulated by petch_trace_neutral_rr[_pair|_batch] (counts_host[0]) and
 * petch_trace_ion_yield[_batch] (counts_host[1]) on this context since the last reset (work accounting
 * for throughput reports; no reference counterpart). counts_host may be NULL; reset != 0 zeroes
 * them after the read. Synchronises when reading. */
int petch_trace_hit_counts(petch_ctx* ctx, int reset, unsigned long long* counts_host);
/* _trace_ff (threed.py:908-920) with the launch of threed.py:949-955: n_ff cosine rays from
 * centroid + 1e-3*n of every face; hitface[j*n_ff + k] = first face hit or -1 (escape). */
int petch_trace_ff(petch_ctx* ctx, const petch_bvh* bvh, const petch_box* box,
                   const float* centroids, const float* gas_normals, int nf, int n_ff,
                   uint64_t seed, int* hitface);

/* ---------------------------------------------------------------- a11-a13, a16: per face - */
/* _norm_clip (threed.py:165-170): out = clamp((fl / max(area,1e-9)) * scale, lo, hi); pass
 * hi = +inf for the unclipped faithful-ion yields (threed.py:814-821). */
int petch_normalize(petch_ctx* ctx, const float* fl, const float* areas, int F, float scale,
                    float lo, float hi, float* out);
/* _cos_k (threed.py:179-185): out = ang/fl where fl > 0, else 0. */
int petch_cos_ratio(petch_ctx* ctx, const float* fl, const float* ang, int F, float* out);
/* smooth_flux (threed.py:129-148, blend_each = 0: n_iter passes then one blend with the input)
 * or smooth_flux_dev applied n_iter times (threed.py:195-214, blend_each = 1). nbr is the table
 * of petch_face_adjacency, normals unit face normals. src != out. */
int petch_smooth(petch_ctx* ctx, const float* src, const int* nbr, const float* normals, int F,
                 int n_iter, float alpha, int blend_each, float* out);
/* Adjoint of petch_normalize at fl: gfl = gout * scale / max(area,1e-9) where
 * lo <= (fl/area) scale <= hi, else 0. No reference counterpart. */
int petch_normalize_adjoint(petch_ctx* ctx, const float* fl, const float* areas, int F,
                            float scale, float lo, float hi, const float* gout, float* gfl);
/* Adjoint of petch_smooth with the same (nbr, normals, n_iter, alpha, blend_each): the passes'
 * transposes in reverse order, gsrc = d loss / d src for gout = d loss / d out (the smoothing is
 * linear). Uses that edge adjacency is symmetric, as on a manifold mesh. gout != gsrc. No
 * reference counterpart. */
int petch_smooth_adjoint(petch_ctx* ctx, const int* nbr, const float* normals, int F, int n_iter,
                         float alpha, int blend_each, const float* gout, float* gsrc);
/* _belen_coverages + the bare update (threed.py:715-737, 903-904): bare = clip(1-thF-thO, 0, 1).
 * With iy_ie/iy_ip (the normalised faithful-ion yields, par['_ion_yield']) GY comes from them,
 * else from m_i, cos_i and the analytic yields. thF/thO may be NULL. */
int petch_belen_coverage(petch_ctx* ctx, const float* m_i, const float* m_F, const float* m_O,
                         const float* cos_i, const float* iy_ie, const float* iy_ip,
                         const petch_chem_params* cp, int F, float* bare, float* thF,
                         float* thO);

/* The end of one neutral fixed-point round of many features in one launch (no reference
 * counterpart: the reference finishes one feature at a time, threed.py:881-904): per feature,
 * m_F and m_O = the F and O accumulators flF and flO normalised by petch_normalize's clip
 * (lo 0, hi) and, with smooth = 1, smoothed by one petch_smooth pass with blend alpha; then
 * bare = petch_belen_coverage of (m_i, m_F, m_O, cos_i, iy_ie, iy_ip). The same arithmetic as
 * those three calls, so the outputs equal theirs. */
typedef struct petch_round_item {
  const float *flF, *flO, *area, *normals;
  const int* nbr;
  const float *m_i, *cos_i, *iy_ie, *iy_ip;
  float *m_F, *m_O, *bare;
  petch_chem_params chem;
  int nf, smooth;
  float scale, hi, alpha;
} petch_round_item;
int petch_finish_round_batch(petch_ctx* ctx, int n, const petch_round_item* items_host);

/* The end of the faithful-ion trace of many features in one launch (no reference counterpart:
 * the reference finishes one feature at a time, threed.py:812-821): per feature, acc is the
 * trace's (5, nf) accumulator (sp, ie, ip, arrivals, angle sum); out (5, nf) receives the three
 * yields and m_i (petch_normalize without a clip, then with smooth = 1 one petch_smooth pass
 * with blend alpha) and cos_i (petch_cos_ratio of arrivals and angle sum). The same arithmetic as
 * those calls, so the outputs equal theirs. */
typedef struct petch_ion_finish_item {
  const float *acc, *area, *normals;
  const int* nbr;
  float* out;
  int nf, smooth;
  float scale, alpha;
} petch_ion_finish_item;
int petch_finish_ion_batch(petch_ctx* ctx, int n, const petch_ion_finish_item* items_host);
/* Closed-form adjoint of petch_belen_coverage for gbare = d loss / d bare (zero where the clip
 * of 1 - thF - thO is active): gm_F, gm_O = d loss / d m_F, d m_O; with iy_ie/iy_ip also
 * g_ie, g_ip = d loss / d iy_ie, d iy_ip, else gm_i = d loss / d m_i (cos_i held constant).
 * gm_i, g_ie and g_ip may be NULL. No reference counterpart (the reference does not
 * differentiate _belen_coverages, threed.py:715-737). */
int petch_belen_coverage_adjoint(petch_ctx* ctx, const float* m_i, const float* m_F,
                                 const float* m_O, const float* cos_i, const float* iy_ie,
                                 const float* iy_ip, const petch_chem_params* cp, int F,
                                 const float* gbare, float* gm_F, float* gm_O, float* gm_i,
                                 float* g_ie, float* g_ip);
/* surface_rate (chemistry.py:83-92 -> belen.py:26-63 or chemistry.py:62-80), V = 0 on mask
 * faces, non-finite V -> 0 (threed.py:1723-1728). Mask faces are is_mask[f] != 0 when is_mask is
 * given, else faces_in_mask of the centroids (threed.py:1608-1616). red[0] = max(V) as float
 * bits (V >= 0), red[1] = number of non-finite values replaced. iy_* may be NULL (analytic
 * yields). */
int petch_surface_rate(petch_ctx* ctx, const float* m_i, const float* m_F, const float* m_O,
                       const float* cos_i, const float* iy_sp, const float* iy_ie,
                       const float* iy_ip, const float* centroids, const uint8_t* is_mask,
                       const petch_mask_geom* mg, const petch_chem_params* cp, int F, float* V,
                       unsigned* red);

/* petch_surface_rate for many surfaces in one launch (the features of a batched step; no
 * reference counterpart): each item's V and reduction pair red[0..1] as that call gives them.
 * With a BVH, the item's traversal guard flag (what petch_bvh_status reads) is copied to
 * *bvh_err in the same pass, so one host read returns every feature's rate reduction and guard. */
typedef struct petch_rate_item {
  const float *m_i, *m_F, *m_O, *cos_i, *iy_sp, *iy_ie, *iy_ip, *centroids;
  const uint8_t* is_mask;
  petch_mask_geom mg;
  petch_chem_params chem;
  int F;
  float* V;
  unsigned* red;
  const petch_bvh* bvh;
  int* bvh_err;
} petch_rate_item;
int petch_surface_rate_batch(petch_ctx* ctx, int n, const petch_rate_item* items_host);
/* Closed-form adjoint of petch_surface_rate for gV = d loss / dV: gm_F, gm_O = d loss / d m_F,
 * d m_O; with ion yields (belen) g_sp, g_ie, g_ip = d loss / d iy_sp, iy_ie, iy_ip, else
 * gm_i = d loss / d m_i. Mask faces and non-finite V pass no gradient; cos_i is held constant.
 * gY (3 x F: Yie, Ysp, Yp rows) receives each face's share of d loss / d the chemistry's yield
 * coefficients (analytic-yield path; zero with ion yields). gm_i, g_sp, g_ie, g_ip and gY may
 * be NULL. No reference counterpart. */
int petch_surface_rate_adjoint(petch_ctx* ctx, const float* m_i, const float* m_F,
                               const float* m_O, const float* cos_i, const float* iy_sp,
                               const float* iy_ie, const float* iy_ip, const float* centroids,
                               const uint8_t* is_mask, const petch_mask_geom* mg,
                               const petch_chem_params* cp, int F, const float* gV, float* gm_F,
                               float* gm_O, float* gm_i, float* g_sp, float* g_ie, float* g_ip,
                               float* gY);

/* ---------------------------------------------------------------- a15, a17 --------------- */
/* extend_velocity_gpu + _extend_kernel (threed.py:1324-1346): Fs = V[nearest face] at nodes
 * with |phi| < band, 0 elsewhere (Fs is the full (nx,ny,nz) grid). */
int petch_extend(petch_ctx* ctx, const petch_bvh* bvh, const float* V, const float* phi, int nx,
                 int ny, int nz, double dx, double band, float* Fs);
/* Adjoint of petch_extend (linear in V): gV (nf, zeroed first) = sum of gFs over the band nodes
 * whose nearest face each face is. gFs != gV. No reference counterpart. */
int petch_extend_adjoint(petch_ctx* ctx, const petch_bvh* bvh, const float* phi, int nx, int ny,
                         int nz, double dx, double band, const float* gFs, float* gV);
/* _nearest_face_cov / gpu_warmstart_bare (threed.py:1619-1639): out[i] = src[nearest face of
 * pts[i]] on the given BVH, `miss` when the mesh is empty. */
int petch_nearest_face_value(petch_ctx* ctx, const petch_bvh* bvh, const float* src,
                             const float* pts, int n, float miss, float* out);

/* petch_extend for many grids, each with its own surface, in one pass (the batched etch loop).
 * No reference counterpart: the reference extends one feature per call. */
typedef struct petch_extend_item {
  const petch_bvh* bvh;
  const float* V;
  const float* phi;
  float* Fs;
  int nx, ny, nz;
  double dx;
  double band;
} petch_extend_item;
int petch_extend_batch(petch_ctx* ctx, int n, const petch_extend_item* items);

/* petch_nearest_face_value for many point sets, each on its own surface, in one launch (the
 * batched etch loop's warm starts). No reference counterpart. */
typedef struct petch_nearest_item {
  const petch_bvh* bvh;
  const float* src;
  const float* pts;
  int n;
  float miss;
  float* out;
} petch_nearest_item;
int petch_nearest_face_value_batch(petch_ctx* ctx, int n, const petch_nearest_item* items);

/* ------------------------------------------- deterministic neutral transports (f4) ------- */
/* knudsen_face_flux (knudsen.py:185-241 with conductance_profile 93-151), the neutral tail of
 * mc_flux_3d_knudsen (threed.py:1149-1208), for two sticking coefficients at once (the F and O
 * species of one coverage round): per-slice gas area and perimeter of phi (z = k dx), the floor
 * faces' (into-gas normal z > 0.7, below sub_top, inside the footprint) sink 0.25 clip(s) A per
 * slice with s = bare beta (or min(bare, mean field bare) with sink_field, threed.py:1193-1197),
 * the conductance chain's tridiagonal solve, and each face's slice concentration (faces at or
 * above sub_top read 1), clipped to [0, 8], into m_a (beta_a) and m_b (beta_b; may be NULL). */
typedef struct petch_knudsen_params {
  double dx, sub_top, Ly;            /* Ly: trench length (a trench's sink is per unit length) */
  int trench;                        /* 1 trench (area = mean open x-width), 0 hole */
  double wall_loss_scale;            /* knudsen_wall_loss_scale */
  int front_loss;                    /* flags.knudsen_front_loss: per-slice loss scale */
  double band_W, pass_frac, ar_pass; /* its knudsen_front_band_W / _passive_frac / _front_ar_pass */
  int footprint;                     /* restrict the sink to |x - cx| (trench) or the radius
                                        (hole) <= half_width + 2 dx; W = 2 half_width */
  double cx, cy, half_width;
  int sink_field;                    /* knudsen_sink = 'field' */
  double beta_a, beta_b;
} petch_knudsen_params;
int petch_knudsen_face_flux(petch_ctx* ctx, const float* phi, int nx, int ny, int nz,
                            const float* centroids, const float* gas_normals, const float* areas,
                            const float* bare, int F, const petch_knudsen_params* p, float* m_a,
                            float* m_b);

/* dda_neutral_flux (dda.py:138-220), the neutrals of mc_flux_3d_dda (threed.py:1211-1244), for
 * two sticking coefficients at once: the march of the _dda_gather_kernel (dda.py:26-95) from
 * every centroid along the 2 n_dir fib_sphere directions, cosine-weighted by the into-gas
 * normal, first from the sky (1 at z >= z_top = (nz - 1.5) dx, 0 at walls), then n_reemit
 * passes with the walls' re-emission field (area-weighted per solid cell, dilated into unseeded
 * solid neighbours), Aitken on the last three iterates, clipped to [0, 8]. Faces' areas are
 * floored at 1e-9. periodic_y wraps the march along y. m_b may be NULL. */
typedef struct petch_dda_params {
  double dx;
  int n_dir, n_reemit, periodic_y;
  double beta_a, beta_b;
} petch_dda_params;
int petch_dda_neutral_flux(petch_ctx* ctx, const float* phi, int nx, int ny, int nz,
                           const float* centroids, const float* gas_normals, const float* areas,
                           const float* bare, int F, const petch_dda_params* p, float* m_a,
                           float* m_b);

/* ---------------------------------------------------------------- a21: radiosity -------- */
/* Form-factor assembly of mc_flux_3d_radiosity (threed.py:956-960) from the petch_trace_ff
 * output: D[j] = escapes of face j / n_ff (float64), and the rays j -> i as a CSR of rows i
 * (row_ptr (nf+1), col up to nf*n_ff entries holding the source face of each ray, rows in ray
 * order). */
int petch_radiosity_assemble(petch_ctx* ctx, const int* hitface, int nf, int n_ff, int* row_ptr,
                             int* col, double* D);
/* Solve G = D + Aff diag(1 - clip(bare*beta,0,1)) G (threed.py:965-985) in float64 by Jacobi
 * from G0 = D: n_iter > 0 sweeps, or n_iter <= 0 until max|dG| <= tol*max|G| (checked every 50
 * sweeps, at most 20000). G (nf, float64) receives the solution, out = clip(G, 0, 8) as float32.
 * *iters_host (may be NULL) = sweeps done. */
int petch_radiosity_solve(petch_ctx* ctx, const int* row_ptr, const int* col, const double* D,
                          int nf, int n_ff, const float* bare, double beta, int n_iter, double tol,
                          double* G, float* out, int* iters_host);
/* Adjoint of petch_radiosity_solve at its solution G (float64) for the output gradient gm
 * (float32, d loss / d out): the transposed system lambda = gG + diag(q) Aff^T lambda, with
 * q = 1 - clip(bare*beta,0,1) and gG = gm where 0 <= G <= 8, solved by Jacobi from lambda0 = gG
 * under the same n_iter / tol rule as the solve; Aff^T is read straight from hitface (the
 * petch_trace_ff output the CSR was assembled from). gD (nf, float64) = lambda = d loss / dD;
 * gbare (nf, float32) = -beta [0 <= bare*beta <= 1] G (Aff^T lambda) = d loss / d bare.
 * gD must not alias G (it holds the sweeps' iterates).
 * No reference counterpart (the reference does not differentiate its radiosity solve,
 * threed.py:965-985); it is the implicit-function adjoint, exact for a converged solve. */
int petch_radiosity_adjoint(petch_ctx* ctx, const int* hitface, int nf, int n_ff, const double* G,
                            const float* bare, double beta, const float* gm, int n_iter,
                            double tol, double* gD, float* gbare, int* iters_host);

/* ---------------------------------------------------------------- profile r(z) ------------- */
/* Smoothed gas area of every z slice of phi: area[k] = dx^2 * sum over the slice's nodes of
 * H_eps(-phi), H_eps(x) = 0 below -eps, 1 above eps, (1 + x/eps + sin(pi x/eps)/pi)/2 between;
 * float64 (nz), a fixed-order reduction (deterministic). The equivalent radius of a hole is
 * sqrt(area / pi), the profile r(z) a profile loss compares with a target. No counterpart in the
 * reference, whose depth metrics (threed.py:1798-1869) are whole-cell column floors. */
int petch_profile_area(petch_ctx* ctx, const float* phi, int nx, int ny, int nz, double dx,
                       double eps, double* area);
/* Adjoint of petch_profile_area: gphi = -garea[k] dx^2 H_eps'(-phi) (float32, phi's shape) for
 * garea = d loss / d area (float64, nz). gphi must not alias phi. */
int petch_profile_area_adjoint(petch_ctx* ctx, const float* phi, int nx, int ny, int nz,
                               double dx, double eps, const double* garea, float* gphi);

#ifdef __cplusplus
}
#endif
#endif /* PETCH_HIP_H */
