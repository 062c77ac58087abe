// Level-set kernels on the (nx,ny,nz) float32 grid, C order (z fastest):
//   - initial signed distance of the masked-substrate CSG          (threed.py:32-59)
//   - first-order upwind Godunov advection + mask re-pin           (threed.py:1262-1307)
//   - Jacobi Godunov-Eikonal ("fsm") narrow-band reinitialisation   (threed.py:1421-1537)
//
// One thread per node, linear index n = (i*ny + j)*nz + k, so a wavefront reads 64 consecutive
// z-nodes and the +-x / +-y neighbours are 64 consecutive words too: every load is coalesced
// and the three neighbour planes are re-read from L2/MALL, not HBM.
#include "common.h"
#include <math.h>

#include <mutex>

namespace {

constexpr int kBS = 256;

// ------------------------------------------------------------------------------ initial SDF
// Distance from (r,z) to the boundary polyline of the half-plane solid
//   {z < s}  U  {s < z < t, r > hw}
// i.e. floor segment (0..hw, s), wall (hw, s..t), mask top (hw..inf, t).
__device__ double csg_distance(double r, double z, double hw, double s, double t) {
  double dr = r - hw;
  double d1 = (r <= hw) ? fabs(z - s) : sqrt(dr * dr + (z - s) * (z - s));
  double d2;
  if (z < s)
    d2 = sqrt(dr * dr + (z - s) * (z - s));
  else if (z > t)
    d2 = sqrt(dr * dr + (z - t) * (z - t));
  else
    d2 = fabs(dr);
  double d3 = (r >= hw) ? fabs(z - t) : sqrt(dr * dr + (z - t) * (z - t));
  return fmin(d1, fmin(d2, d3));
}

__global__ void k_init_sdf(float* __restrict__ phi, uint8_t* __restrict__ mask, int nx, int ny,
                           int nz, double dx, double cx, double cy, double hw, double hw2,
                           double s, double t, int hole) {
  // every product is rounded on its own (no fused multiply-add), as numpy evaluates it, so the
  // mask decisions at exact ties (|x - Lx/2| == w/2) come out the same
#pragma clang fp contract(off)
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long N = (long long)nx * ny * nz;
  if (n >= N) return;
  int k = (int)(n % nz);
  int j = (int)((n / nz) % ny);
  int i = (int)(n / ((long long)nz * ny));
  // xs = arange(n)*dx evaluated in float64 without contraction, as numpy does
  double X = (double)i * dx;
  double Y = (double)j * dx;
  double Z = (double)k * dx;
  double ax = X - cx;
  bool opening;
  double r;
  if (hole) {
    double ay = Y - cy;
    double r2 = ax * ax + ay * ay;
    opening = r2 < hw2;
    r = sqrt(r2);
  } else {
    r = fabs(ax);
    opening = r < hw;
  }
  bool band = (Z >= s) && (Z < t);
  mask[n] = (band && !opening) ? 1 : 0;
  double d = csg_distance(r, Z, hw, s, t);
  // the junction plane z = s under the mask is interior: substrate and mask are one solid
  bool solid = (Z < s) || (Z >= s && Z < t && r > hw);
  phi[n] = (float)(solid ? d : -d);
}

// ------------------------------------------------------------------------------ advection
__global__ void k_advect(const float* __restrict__ phi, const float* __restrict__ F,
                         const uint8_t* __restrict__ mask, const float* __restrict__ mask_phi,
                         float* __restrict__ out, int nx, int ny, int nz, float inv_dx,
                         float dt) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long N = (long long)nx * ny * nz;
  if (n >= N) return;
  if (mask[n]) {
    out[n] = mask_phi[n];
    return;
  }
  int k = (int)(n % nz);
  int j = (int)((n / nz) % ny);
  int i = (int)(n / ((long long)nz * ny));
  long long sx = (long long)ny * nz, sy = nz;
  float c = phi[n];
  float g = 0.f, dm, dp;
  dm = (i >= 1) ? (c - phi[n - sx]) * inv_dx : 0.f;
  dp = (i <= nx - 2) ? (phi[n + sx] - c) * inv_dx : 0.f;
  g += fmaxf(dm, 0.f) * fmaxf(dm, 0.f) + fminf(dp, 0.f) * fminf(dp, 0.f);
  dm = (j >= 1) ? (c - phi[n - sy]) * inv_dx : 0.f;
  dp = (j <= ny - 2) ? (phi[n + sy] - c) * inv_dx : 0.f;
  g += fmaxf(dm, 0.f) * fmaxf(dm, 0.f) + fminf(dp, 0.f) * fminf(dp, 0.f);
  dm = (k >= 1) ? (c - phi[n - 1]) * inv_dx : 0.f;
  dp = (k <= nz - 2) ? (phi[n + 1] - c) * inv_dx : 0.f;
  g += fmaxf(dm, 0.f) * fmaxf(dm, 0.f) + fminf(dp, 0.f) * fminf(dp, 0.f);
  out[n] = c - dt * F[n] * sqrtf(g);
}

// ------------------------------------------------------------------------------ advection adjoint
// Adjoint of one substep at a state phi: out = phi - dt F sqrt(G) off the mask, mask_phi on it,
// with G = sum over axes of max(D-,0)^2 + min(D+,0)^2. Node m gathers what it contributes to
// its own output and to its six neighbours' G; where G = 0 the sqrt term has no derivative (0
// is taken); mask nodes pass no gradient. gF accumulates (the speed is shared by the substeps).
struct AdjStencil {
  float G, mp[3], mn[3];  // G and, per axis, max(D-,0) and min(D+,0)
};

__device__ __forceinline__ AdjStencil adj_stencil(const float* __restrict__ phi, long long n,
                                                  int i, int j, int k, int nx, int ny, int nz,
                                                  float inv_dx) {
  const long long sx = (long long)ny * nz, sy = nz;
  const float c = phi[n];
  AdjStencil a;
  float dm, dp;
  dm = (i >= 1) ? (c - phi[n - sx]) * inv_dx : 0.f;
  dp = (i <= nx - 2) ? (phi[n + sx] - c) * inv_dx : 0.f;
  a.mp[0] = fmaxf(dm, 0.f);
  a.mn[0] = fminf(dp, 0.f);
  dm = (j >= 1) ? (c - phi[n - sy]) * inv_dx : 0.f;
  dp = (j <= ny - 2) ? (phi[n + sy] - c) * inv_dx : 0.f;
  a.mp[1] = fmaxf(dm, 0.f);
  a.mn[1] = fminf(dp, 0.f);
  dm = (k >= 1) ? (c - phi[n - 1]) * inv_dx : 0.f;
  dp = (k <= nz - 2) ? (phi[n + 1] - c) * inv_dx : 0.f;
  a.mp[2] = fmaxf(dm, 0.f);
  a.mn[2] = fminf(dp, 0.f);
  a.G = a.mp[0] * a.mp[0] + a.mn[0] * a.mn[0] + a.mp[1] * a.mp[1] + a.mn[1] * a.mn[1] +
        a.mp[2] * a.mp[2] + a.mn[2] * a.mn[2];
  return a;
}

// dL/dG at node n: -dt F gout / (2 sqrt G), 0 on the mask or where G = 0
__device__ __forceinline__ float adj_weight(const AdjStencil& a, const float* __restrict__ F,
                                            const uint8_t* __restrict__ mask,
                                            const float* __restrict__ gout, long long n,
                                            float dt) {
  if (mask[n] || !(a.G > 0.f)) return 0.f;
  return -dt * F[n] * gout[n] / (2.f * sqrtf(a.G));
}

// Pass 1 of the adjoint, per node m: w[m] = dL/dG at m (adj_weight of m's own stencil), gF[m]
// += -dt sqrt(G_m) gout[m] off the mask, and gphi[m] = the local terms (the pass-through of
// gout off the mask plus m's own stencil differentiated in phi_m).
__global__ void k_advect_adjoint_local(const float* __restrict__ phi,
                                       const float* __restrict__ F,
                                       const uint8_t* __restrict__ mask,
                                       const float* __restrict__ gout, float* __restrict__ w,
                                       float* __restrict__ gphi, float* __restrict__ gF, int nx,
                                       int ny, int nz, float inv_dx, float dt) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)nx * ny * nz;
  if (m >= N) return;
  const int k = (int)(m % nz);
  const int j = (int)((m / nz) % ny);
  const int i = (int)(m / ((long long)nz * ny));
  const AdjStencil self = adj_stencil(phi, m, i, j, k, nx, ny, nz, inv_dx);
  const float wm = adj_weight(self, F, mask, gout, m, dt);
  const bool msk = mask[m] != 0;
  float g = msk ? 0.f : gout[m];
  for (int ax = 0; ax < 3; ++ax) g += wm * 2.f * (self.mp[ax] - self.mn[ax]) * inv_dx;
  w[m] = wm;
  gphi[m] = g;
  if (!msk) gF[m] += -dt * sqrtf(self.G) * gout[m];
}

// Pass 2, per node m: add what m's six neighbours' stencils owe phi_m. Neighbour n = m + e_ax
// uses phi_m in its D- = (phi_n - phi_m)/dx, neighbour n = m - e_ax in its D+ =
// (phi_m - phi_n)/dx, so each term needs only w[n] and the two values, not n's whole stencil.
__global__ void k_advect_adjoint_gather(const float* __restrict__ phi,
                                        const float* __restrict__ w, float* __restrict__ gphi,
                                        int nx, int ny, int nz, float inv_dx) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)nx * ny * nz;
  if (m >= N) return;
  const int k = (int)(m % nz);
  const int j = (int)((m / nz) % ny);
  const int i = (int)(m / ((long long)nz * ny));
  const long long st[3] = {(long long)ny * nz, (long long)nz, 1};
  const int idx[3] = {i, j, k}, lim[3] = {nx, ny, nz};
  const float c = phi[m];
  float g = gphi[m];
  for (int ax = 0; ax < 3; ++ax) {
    if (idx[ax] + 1 < lim[ax]) {
      const long long n = m + st[ax];
      g -= w[n] * 2.f * fmaxf((phi[n] - c) * inv_dx, 0.f) * inv_dx;
    }
    if (idx[ax] >= 1) {
      const long long n = m - st[ax];
      g += w[n] * 2.f * fminf((c - phi[n]) * inv_dx, 0.f) * inv_dx;
    }
  }
  gphi[m] = g;
}

// ------------------------------------------------------------------------------ FSM reinit
// Exact Godunov update of sum_d (x - a_d)_+^2 = dx^2 with the staged 1 -> 2 -> 3 term form,
// evaluated in float64 relative to the smallest neighbour (y = x - a1) so the three-term root
// has no cancellation against the O(band) magnitudes; the result is rounded to float32 once.
__device__ __forceinline__ float eik_solve(float f1, float f2, float f3, double dx) {
  double a1 = f1, a2 = f2, a3 = f3, t;
  if (a1 > a2) { t = a1; a1 = a2; a2 = t; }
  if (a2 > a3) { t = a2; a2 = a3; a3 = t; }
  if (a1 > a2) { t = a1; a1 = a2; a2 = t; }
  double x = a1 + dx;
  if (x <= a2) return (float)x;
  double b2 = a2 - a1, b3 = a3 - a1;
  double d2 = fmax(2.0 * dx * dx - b2 * b2, 0.0);
  double y = 0.5 * (b2 + sqrt(d2));
  if (a1 + y <= a3) return (float)(a1 + y);
  double s1 = b2 + b3;
  double d3 = fmax(s1 * s1 - 3.0 * (b2 * b2 + b3 * b3 - dx * dx), 0.0);
  return (float)(a1 + (s1 + sqrt(d3)) / 3.0);
}

// Frozen interface seed: nodes with a sign change to a face neighbour get the sub-cell distance
// |phi0| / max(|grad phi0|_central, 0.5), capped at 1.8 dx; every other node starts at `big`.
// Evaluated in float64 from the float32 inputs, rounded once.
__global__ void k_eik_init(const float* __restrict__ phi0, float* __restrict__ U,
                           uint8_t* __restrict__ frozen, int nx, int ny, int nz, double dx,
                           float big) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long N = (long long)nx * ny * nz;
  if (n >= N) return;
  int k = (int)(n % nz);
  int j = (int)((n / nz) % ny);
  int i = (int)(n / ((long long)nz * ny));
  long long sx = (long long)ny * nz, sy = nz;
  float c = phi0[n];
  float oxm = phi0[i > 0 ? n - sx : n], oxp = phi0[i < nx - 1 ? n + sx : n];
  float oym = phi0[j > 0 ? n - sy : n], oyp = phi0[j < ny - 1 ? n + sy : n];
  float ozm = phi0[k > 0 ? n - 1 : n], ozp = phi0[k < nz - 1 ? n + 1 : n];
  // sign change to a face neighbour (products of floats are exact in double)
  double cd = c;
  bool iface = (cd * oxm < 0.0) || (cd * oxp < 0.0) || (cd * oym < 0.0) || (cd * oyp < 0.0) ||
               (cd * ozm < 0.0) || (cd * ozp < 0.0);
  if (iface) {
    double h = 0.5 / dx;
    double gx = ((double)oxp - oxm) * h, gy = ((double)oyp - oym) * h, gz = ((double)ozp - ozm) * h;
    double grad0 = fmax(sqrt(gx * gx + gy * gy + gz * gz), 0.5);
    U[n] = (float)fmin(fabs(cd) / grad0, 1.8 * dx);
    frozen[n] = 1;
  } else {
    U[n] = big;
    frozen[n] = 0;
  }
}

__global__ void k_eik_jacobi(const float* __restrict__ U, const uint8_t* __restrict__ frozen,
                             float* __restrict__ out, int nx, int ny, int nz, double dx,
                             float big) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long N = (long long)nx * ny * nz;
  if (n >= N) return;
  float u = U[n];
  if (frozen[n]) {
    out[n] = u;
    return;
  }
  int k = (int)(n % nz);
  int j = (int)((n / nz) % ny);
  int i = (int)(n / ((long long)nz * ny));
  long long sx = (long long)ny * nz, sy = nz;
  float ax = fminf(i >= 1 ? U[n - sx] : big, i <= nx - 2 ? U[n + sx] : big);
  float ay = fminf(j >= 1 ? U[n - sy] : big, j <= ny - 2 ? U[n + sy] : big);
  float az = fminf(k >= 1 ? U[n - 1] : big, k <= nz - 2 ? U[n + 1] : big);
  out[n] = fminf(u, eik_solve(ax, ay, az, dx));
}

// phi <- sign(phi0) * min(U, clampv), sign(0) = +1, NaN propagates.
__global__ void k_eik_final(float* __restrict__ phi, const float* __restrict__ U, long long N,
                            float clampv) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float p = phi[n];
  float m = fminf(U[n], clampv);
  float s = (p < 0.f) ? -1.f : 1.f;
  phi[n] = (p != p) ? p : s * m;
}

}  // namespace

extern "C" int petch_init_sdf(petch_ctx* ctx, float* phi, uint8_t* mask, int nx, int ny, int nz,
                              double dx, double Lx, double Ly, double trench_width,
                              double mask_th, double sub_top, int hole) {
  PETCH_ARG(ctx, ctx && phi && mask, "init_sdf: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "init_sdf: bad shape %d,%d,%d", nx, ny, nz);
  PETCH_ARG(ctx, dx > 0, "init_sdf: dx must be > 0");
  long long N = (long long)nx * ny * nz;
  double hw = trench_width / 2.0;
  double hw2 = hw * hw;
  double t = sub_top + mask_th;
  hipLaunchKernelGGL(k_init_sdf, dim3(petch_blocks(N, kBS)), dim3(kBS), 0, ctx->stream, phi, mask,
                     nx, ny, nz, dx, Lx / 2.0, Ly / 2.0, hw, hw2, sub_top, t, hole);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_advect(petch_ctx* ctx, float* phi, float* work, const float* F,
                            const uint8_t* mask, const float* mask_phi, int nx, int ny, int nz,
                            double dx, double dt, int nsub) {
  PETCH_ARG(ctx, ctx && phi && work && F && mask && mask_phi, "advect: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "advect: bad shape %d,%d,%d", nx, ny, nz);
  PETCH_ARG(ctx, nsub >= 0, "advect: nsub must be >= 0");
  long long N = (long long)nx * ny * nz;
  float inv = (float)(1.0 / dx);
  float fdt = (float)dt;
  float* a = phi;
  float* b = work;
  for (int s = 0; s < nsub; ++s) {
    hipLaunchKernelGGL(k_advect, dim3(petch_blocks(N, kBS)), dim3(kBS), 0, ctx->stream, a, F,
                       mask, mask_phi, b, nx, ny, nz, inv, fdt);
    PETCH_LAUNCH_CHECK(ctx);
    float* tmp = a; a = b; b = tmp;
  }
  if (a != phi)
    PETCH_HIP(ctx, hipMemcpyAsync(phi, a, N * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
  return PETCH_OK;
}

extern "C" int petch_advect_step(petch_ctx* ctx, const float* phi, const float* F,
                                 const uint8_t* mask, const float* mask_phi, int nx, int ny,
                                 int nz, double dx, double dt, float* out) {
  PETCH_ARG(ctx, ctx && phi && F && mask && mask_phi && out, "advect_step: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "advect_step: bad shape %d,%d,%d", nx, ny, nz);
  PETCH_ARG(ctx, out != phi, "advect_step: out must not alias phi");
  const long long N = (long long)nx * ny * nz;
  hipLaunchKernelGGL(k_advect, dim3(petch_blocks(N, kBS)), dim3(kBS), 0, ctx->stream, phi, F,
                     mask, mask_phi, out, nx, ny, nz, (float)(1.0 / dx), (float)dt);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

namespace {

// The reinit's launches over a given workspace: U and U2 (N floats each, ping-pong) and the
// frozen flags (N bytes). k_eik_init writes every node of U and of the flags before any read,
// so the workspace needs no clearing.
int reinit_launch(petch_ctx* ctx, float* phi, int nx, int ny, int nz, double dx, double band,
                  int n_iter, float* U, float* U2, uint8_t* fr) {
  const long long N = (long long)nx * ny * nz;
  const double big = band + 4.0 * dx;
  if (n_iter < 0) n_iter = (int)ceil(band / dx) + 4;
  dim3 g(petch_blocks(N, kBS)), b(kBS);
  hipLaunchKernelGGL(k_eik_init, g, b, 0, ctx->stream, phi, U, fr, nx, ny, nz, dx, (float)big);
  PETCH_LAUNCH_CHECK(ctx);
  for (int it = 0; it < n_iter; ++it) {
    hipLaunchKernelGGL(k_eik_jacobi, g, b, 0, ctx->stream, U, fr, U2, nx, ny, nz, dx, (float)big);
    PETCH_LAUNCH_CHECK(ctx);
    float* t = U; U = U2; U2 = t;
  }
  hipLaunchKernelGGL(k_eik_final, g, b, 0, ctx->stream, phi, U, N, (float)(band + dx));
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of the reinit workspace for N nodes: U, U2 and the flags, each 256-byte aligned.
size_t reinit_workspace_bytes(long long N) {
  return 2 * round256((size_t)N * sizeof(float)) + round256((size_t)N);
}

constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
struct NoAllocSlot {
  int state;        // 0 not resolved, 1 resolved, -1 failed
  void* ptr;
  hipError_t err;
};
NoAllocSlot g_noalloc[kMaxDevices];
std::mutex g_noalloc_mu;

}  // namespace

// Fixed workspace of petch_reinit_fsm_noalloc: one copy per GPU, allocated when the library's
// code object is loaded on that GPU and held until the process ends. Host code never uses this
// name as an address: its device address comes from hipGetSymbolAddress, per device, at context
// creation (petch_noalloc_setup) and is kept in the context.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Resolves the device address of this GPU's copy of the fixed workspace, once per device, with
// that device current (petch_ctx_create calls it right after hipSetDevice). A failure is kept
// and reported by petch_reinit_fsm_noalloc; the other entry points do not need the workspace.
void petch_noalloc_setup(petch_ctx* ctx) {
  std::lock_guard<std::mutex> guard(g_noalloc_mu);
  ctx->noalloc_scratch = nullptr;
  ctx->noalloc_err = hipErrorInvalidDevice;
  if (ctx->device >= 0 && ctx->device < kMaxDevices) {
    NoAllocSlot& s = g_noalloc[ctx->device];
    if (s.state == 0) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      s.state = (e == hipSuccess && p) ? 1 : -1;
      s.ptr = (e == hipSuccess) ? p : nullptr;
      s.err = (e == hipSuccess && !p) ? hipErrorInvalidSymbol : e;
    }
    ctx->noalloc_scratch = s.ptr;
    ctx->noalloc_err = s.err;
  }
}

extern "C" int petch_reinit_fsm(petch_ctx* ctx, float* phi, int nx, int ny, int nz, double dx,
                                double band, int n_iter) {
  PETCH_ARG(ctx, ctx && phi, "reinit_fsm: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "reinit_fsm: bad shape %d,%d,%d", nx, ny, nz);
  PETCH_ARG(ctx, dx > 0 && band > 0, "reinit_fsm: dx and band must be > 0");
  long long N = (long long)nx * ny * nz;
  float* U = (float*)petch_scratch(ctx, SCR_LS_A, N * sizeof(float));
  float* U2 = (float*)petch_scratch(ctx, SCR_LS_B, N * sizeof(float));
  uint8_t* fr = (uint8_t*)petch_scratch(ctx, SCR_LS_FROZEN, N);
  if (!U || !U2 || !fr) return PETCH_E_HIP;
  return reinit_launch(ctx, phi, nx, ny, nz, dx, band, n_iter, U, U2, fr);
}

extern "C" int petch_reinit_fsm_noalloc(petch_ctx* ctx, float* phi, int nx, int ny, int nz,
                                        double dx, double band, int n_iter) {
  PETCH_ARG(ctx, ctx && phi, "reinit_fsm_noalloc: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "reinit_fsm_noalloc: bad shape %d,%d,%d", nx, ny,
            nz);
  PETCH_ARG(ctx, dx > 0 && band > 0, "reinit_fsm_noalloc: dx and band must be > 0");
  const long long N = (long long)nx * ny * nz;
  const size_t need = reinit_workspace_bytes(N);
  if (need > kNoAllocScratchBytes)
    return petch_set_error(ctx, PETCH_E_CAPACITY,
                           "reinit_fsm_noalloc: a %dx%dx%d grid needs %zu workspace bytes; the "
                           "fixed workspace holds %zu (grids up to about %lld nodes); use "
                           "petch_reinit_fsm",
                           nx, ny, nz, need, kNoAllocScratchBytes,
                           (long long)(kNoAllocScratchBytes / 9));
  if (!ctx->noalloc_scratch)
    return petch_set_error(ctx, PETCH_E_HIP,
                           "reinit_fsm_noalloc: the fixed workspace of device %d is unavailable "
                           "(hipGetSymbolAddress: %s)",
                           ctx->device, hipGetErrorString(ctx->noalloc_err));
  unsigned char* base = (unsigned char*)ctx->noalloc_scratch;
  float* U = (float*)base;
  float* U2 = (float*)(base + round256((size_t)N * sizeof(float)));
  uint8_t* fr = base + 2 * round256((size_t)N * sizeof(float));
  return reinit_launch(ctx, phi, nx, ny, nz, dx, band, n_iter, U, U2, fr);
}

extern "C" int petch_advect_adjoint(petch_ctx* ctx, const float* phi, const float* F,
                                    const uint8_t* mask, const float* gout, int nx, int ny, int nz,
                                    double dx, double dt, float* gphi, float* gF) {
  PETCH_ARG(ctx, ctx && phi && F && mask && gout && gphi && gF, "advect_adjoint: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "advect_adjoint: bad shape %d,%d,%d", nx, ny, nz);
  PETCH_ARG(ctx, gphi != gout && gphi != phi && gphi != gF,
            "advect_adjoint: gphi must not alias gout, phi or gF");
  const long long N = (long long)nx * ny * nz;
  float* w = (float*)petch_scratch(ctx, SCR_LS_A, N * sizeof(float));
  if (!w) return PETCH_E_HIP;
  const float inv = (float)(1.0 / dx);
  dim3 g(petch_blocks(N, kBS)), b(kBS);
  hipLaunchKernelGGL(k_advect_adjoint_local, g, b, 0, ctx->stream, phi, F, mask, gout, w, gphi, gF,
                     nx, ny, nz, inv, (float)dt);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_advect_adjoint_gather, g, b, 0, ctx->stream, phi, w, gphi, nx, ny, nz, inv);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}
