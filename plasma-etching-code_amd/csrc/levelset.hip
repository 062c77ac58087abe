// Level-set kernels on the (nx,ny,nz) float32 grid, C order (z fastest):
//   - initial signed distance of the masked-substrate CSG          (threed.py:32-59)
//   - first-order upwind Godunov advection + mask re-pin           (threed.py:1262-1307)
//   - Jacobi Godunov-Eikonal ("fsm") narrow-band reinitialisation   (threed.py:1421-1537)
//
// One thread per node, linear index n = (i*ny + j)*nz + k, so a wavefront reads 64 consecutive
// z-nodes and the +-x / +-y neighbours are 64 consecutive words too: every load is coalesced
// and the three neighbour planes are re-read from L2/MALL, not HBM.
#include "common.h"
#include "fastdiv.h"
#include <math.h>
#include <stdlib.h>

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

// Node n of an (nx, ny, nz) grid as (i, j, k), by 32-bit division by the invariant ny and nz
// (node indices of one grid fit 32 bits; the entry points check).
__device__ __forceinline__ void split_node(long long n, const petch::FastDiv& fy,
                                           const petch::FastDiv& fz, int& i, int& j, int& k) {
  const uint32_t u = (uint32_t)n;
  const uint32_t c = petch::fastdiv(u, fz);
  k = (int)(u - c * fz.d);
  i = (int)petch::fastdiv(c, fy);
  j = (int)(c - (uint32_t)i * fy.d);
}

// ------------------------------------------------------------------------------ advection
// One upwind substep at node n = (i, j, k) of an (nx, ny, nz) grid.
__device__ __forceinline__ void advect_node(const float* __restrict__ phi,
                                            const float* __restrict__ F,
                                            const uint8_t* __restrict__ mask,
                                            const float* __restrict__ mask_phi,
                                            float* __restrict__ out, int nx, int ny, int nz,
                                            float inv_dx, float dt, long long n, int i, int j,
                                            int k) {
  if (mask[n]) {
    out[n] = mask_phi[n];
    return;
  }
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

__global__ void k_advect(const float* __restrict__ phi, const float* __restrict__ F,
                         const uint8_t* __restrict__ mask, const float* __restrict__ mask_phi,
                         float* __restrict__ out, int nx, int ny, int nz, float inv_dx,
                         float dt, petch::FastDiv fy, petch::FastDiv fz) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long N = (long long)nx * ny * nz;
  if (n >= N) return;
  int i, j, k;
  split_node(n, fy, fz, i, j, k);
  advect_node(phi, F, mask, mask_phi, out, nx, ny, nz, inv_dx, dt, n, i, j, k);
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
__device__ __forceinline__ void eik_init_node(const float* __restrict__ phi0,
                                              float* __restrict__ U,
                                              uint8_t* __restrict__ frozen, int nx, int ny,
                                              int nz, double dx, float big, long long n, int i,
                                              int j, int k) {
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

__global__ void k_eik_init(const float* __restrict__ phi0, float* __restrict__ U,
                           uint8_t* __restrict__ frozen, int nx, int ny, int nz, double dx,
                           float big, petch::FastDiv fy, petch::FastDiv fz) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long N = (long long)nx * ny * nz;
  if (n >= N) return;
  int i, j, k;
  split_node(n, fy, fz, i, j, k);
  eik_init_node(phi0, U, frozen, nx, ny, nz, dx, big, n, i, j, k);
}

__device__ __forceinline__ void eik_jacobi_node(const float* __restrict__ U,
                                                const uint8_t* __restrict__ frozen,
                                                float* __restrict__ out, int nx, int ny, int nz,
                                                double dx, float big, long long n, int i,
                                                int j, int k) {
  float u = U[n];
  if (frozen[n]) {
    out[n] = u;
    return;
  }
  long long sx = (long long)ny * nz, sy = nz;
  float ax = fminf(i >= 1 ? U[n - sx] : big, i <= nx - 2 ? U[n + sx] : big);
  float ay = fminf(j >= 1 ? U[n - sy] : big, j <= ny - 2 ? U[n + sy] : big);
  float az = fminf(k >= 1 ? U[n - 1] : big, k <= nz - 2 ? U[n + 1] : big);
  out[n] = fminf(u, eik_solve(ax, ay, az, dx));
}

__global__ void k_eik_jacobi(const float* __restrict__ U, const uint8_t* __restrict__ frozen,
                             float* __restrict__ out, int nx, int ny, int nz, double dx,
                             float big, petch::FastDiv fy, petch::FastDiv fz) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long N = (long long)nx * ny * nz;
  if (n >= N) return;
  int i, j, k;
  split_node(n, fy, fz, i, j, k);
  eik_jacobi_node(U, frozen, out, nx, ny, nz, dx, big, n, i, j, k);
}

// phi <- sign(phi0) * min(U, clampv), sign(0) = +1, NaN propagates.
__device__ __forceinline__ void eik_final_node(float* __restrict__ phi,
                                               const float* __restrict__ U, float clampv,
                                               long long n) {
  float p = phi[n];
  float m = fminf(U[n], clampv);
  float s = (p < 0.f) ? -1.f : 1.f;
  phi[n] = (p != p) ? p : s * m;
}

__global__ void k_eik_final(float* __restrict__ phi, const float* __restrict__ U, long long N,
                            float clampv) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  eik_final_node(phi, U, clampv, n);
}

// One step of the PDE reinit of reinit_method='gpu' (threed.py:1349-1398), float32 in the
// reference's order: interface nodes (phi0 changes sign to a face neighbour) relax toward
// D = clip(phi0 / max(|grad phi0|, 0.5), +-1.8 dx); the others take the Godunov upwind
// |grad phi| with the smoothed sign phi0 / sqrt(phi0^2 + (dx/2)^2). Neighbours clamp at the
// boundary.
__global__ void k_reinit_pde(const float* __restrict__ phi, const float* __restrict__ phi0,
                             float* __restrict__ out, int nx, int ny, int nz, float inv_dx,
                             float dtau, petch::FastDiv fy, petch::FastDiv fz) {
#pragma clang fp contract(off)
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long long)nx * ny * nz) return;
  int i, j, k;
  split_node(n, fy, fz, i, j, k);
  const long long sx = (long long)ny * nz, sy = nz;
  const long long xm = i > 0 ? n - sx : n, xp = i < nx - 1 ? n + sx : n;
  const long long ym = j > 0 ? n - sy : n, yp = j < ny - 1 ? n + sy : n;
  const long long zm = k > 0 ? n - 1 : n, zp = k < nz - 1 ? n + 1 : n;
  const float c = phi[n];
  const float s0 = phi0[n];
  const float o_xm = phi0[xm], o_xp = phi0[xp], o_ym = phi0[ym], o_yp = phi0[yp];
  const float o_zm = phi0[zm], o_zp = phi0[zp];
  const bool iface = (s0 * o_xm < 0.f) || (s0 * o_xp < 0.f) || (s0 * o_ym < 0.f) ||
                     (s0 * o_yp < 0.f) || (s0 * o_zm < 0.f) || (s0 * o_zp < 0.f);
  if (iface) {
    const float gx0 = (o_xp - o_xm) * 0.5f * inv_dx;
    const float gy0 = (o_yp - o_ym) * 0.5f * inv_dx;
    const float gz0 = (o_zp - o_zm) * 0.5f * inv_dx;
    const float grad0 = fmaxf(sqrtf(gx0 * gx0 + gy0 * gy0 + gz0 * gz0), 0.5f);
    const float dxl = 1.0f / inv_dx;
    const float lim = 1.8f * dxl;
    const float D = fminf(fmaxf(s0 / grad0, -lim), lim);
    const float sgn0 = s0 < 0.f ? -1.f : 1.f;
    out[n] = c - dtau * inv_dx * (sgn0 * fabsf(c) - D);
    return;
  }
  const float Dxm = (c - phi[xm]) * inv_dx, Dxp = (phi[xp] - c) * inv_dx;
  const float Dym = (c - phi[ym]) * inv_dx, Dyp = (phi[yp] - c) * inv_dx;
  const float Dzm = (c - phi[zm]) * inv_dx, Dzp = (phi[zp] - c) * inv_dx;
  const float sgn = s0 / sqrtf(s0 * s0 + 0.25f / (inv_dx * inv_dx));
  float gx, gy, gz;
  if (sgn > 0.f) {
    gx = fmaxf(fmaxf(Dxm, 0.f), -fminf(Dxp, 0.f));
    gy = fmaxf(fmaxf(Dym, 0.f), -fminf(Dyp, 0.f));
    gz = fmaxf(fmaxf(Dzm, 0.f), -fminf(Dzp, 0.f));
  } else {
    gx = fmaxf(-fminf(Dxm, 0.f), fmaxf(Dxp, 0.f));
    gy = fmaxf(-fminf(Dym, 0.f), fmaxf(Dyp, 0.f));
    gz = fmaxf(-fminf(Dzm, 0.f), fmaxf(Dzp, 0.f));
  }
  const float G = sqrtf(gx * gx + gy * gy + gz * gz);
  out[n] = c - dtau * sgn * (G - 1.f);
}

// One axis pass of the separable Gaussian of surf_smooth: out = sum over |j| <= r of
// w[r + j] in[reflect(i + j)] along `axis`, in float64 as w[r] x0 + sum_{j=r..1} w[r+j]
// (x[-j] + x[+j]), stored as float32; 'reflect' boundaries (d c b a | a b c d | d c b a).
// With mask, nodes with mask[n] take mask_phi[n] (the last pass re-pins the mask).
__global__ void k_gauss_axis(const float* __restrict__ in, float* __restrict__ out, int nx,
                             int ny, int nz, int axis, const double* __restrict__ w, int r,
                             petch::FastDiv fy, petch::FastDiv fz, const uint8_t* __restrict__ mask,
                             const float* __restrict__ mask_phi) {
#pragma clang fp contract(off)
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)nx * ny * nz;
  if (n >= N) return;
  if (mask && mask[n]) {
    out[n] = mask_phi[n];
    return;
  }
  int i, j, k;
  split_node(n, fy, fz, i, j, k);
  const int len = axis == 0 ? nx : (axis == 1 ? ny : nz);
  const int pos = axis == 0 ? i : (axis == 1 ? j : k);
  const long long stride = axis == 0 ? (long long)ny * nz : (axis == 1 ? (long long)nz : 1LL);
  const long long base = n - (long long)pos * stride;
  double acc = w[r] * (double)in[n];
  for (int q = r; q >= 1; --q) {
    int a = pos - q, b = pos + q;
    while (a < 0 || a >= len) a = a < 0 ? -a - 1 : 2 * len - a - 1;
    while (b < 0 || b >= len) b = b < 0 ? -b - 1 : 2 * len - b - 1;
    acc = acc + ((double)in[base + (long long)a * stride] + (double)in[base + (long long)b * stride]) * w[r + q];
  }
  out[n] = (float)acc;
}

// ------------------------------------------------------------------------------ many grids
// One launch for the same stage of many grids (the features of a batched step). Full passes map
// workgroups to nodes: grid f owns the workgroups [nblk0[f], nblk0[f+1]) of the launch, found by
// a binary search that is uniform across the workgroup, and its workgroup covers `rounds` runs
// of kBS consecutive nodes (node (offset x rounds + r) x kBS + lane in round r). Up to 8 rounds
// keep a full pass over ~100M nodes near 50k workgroups: with one run per workgroup its ~400k
// workgroups made the pass bound by workgroup dispatch (0.4-0.75 ms against ~0.1 ms of
// memory traffic). Stages that leave most of a grid unchanged
// (advection substeps, Jacobi sweeps) work on bricks instead: a list of (grid, brick) entries is
// built on the device for the launch, and a fixed number of workgroups loops over it, so the
// bricks left out cost neither a workgroup dispatch nor a memory access.
struct LsGrid {
  float* phi;
  float* work;               // advection ping-pong partner of phi
  const float* F;
  const uint8_t* mask;
  const float* mask_phi;
  float *U, *U2;             // reinit workspace slices
  uint8_t* fr;
  uint8_t* bfl;              // per-brick activity flags (one plane, or count + 1 for the reinit)
  int nx, ny, nz;
  int bnx, bny, bnz;         // bricks per axis
  petch::FastDiv fy, fz, fby, fbz;   // division by ny, nz, bny, bnz
  int count;                 // advection substeps or Jacobi sweeps
  double dx;
  float inv_dx, dt, big, clampv;
  long long nblk0;           // first workgroup of this grid in a full pass
  int rounds;                // nodes of a full-pass workgroup: rounds x kBS consecutive ones
  long long bblk0;           // first brick of this grid in the launch's brick numbering
};

// Bricks are 4 x 4 x 16 nodes (z fastest inside, bk fastest across); in a brick launch lane t
// takes node (4 bi + t / 64, 4 bj + (t / 16) % 4, 16 bk + t % 16), so each 16-lane group reads a
// 64-byte run of one z column. A node's six neighbours lie in its brick or a face neighbour.
constexpr int kBX = 4, kBY = 4, kBZ = 16;
static_assert(kBX * kBY * kBZ == kBS, "one brick per workgroup pass");
constexpr int kListWG = 2048;   // workgroups of a brick-list launch (8 per CU)
constexpr int kHalo = 0x40000000;  // list entry flag: copy the brick, do not update it

__device__ __forceinline__ int grid_of_nodes(const LsGrid* __restrict__ g, int n) {
  int lo = 0, hi = n - 1;
  const long long b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g[mid].nblk0 <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int grid_of_brick(const LsGrid* __restrict__ g, int n, long long gb) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g[mid].bblk0 <= gb) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ long long nbricks(const LsGrid& G) {
  return (long long)G.bnx * G.bny * G.bnz;
}

__device__ __forceinline__ long long brick_of_node(const LsGrid& G, long long n) {
  int i, j, k;
  split_node(n, G.fy, G.fz, i, j, k);
  return ((long long)(i / kBX) * G.bny + j / kBY) * G.bnz + k / kBZ;
}

// brick b as (bi, bj, bk)
__device__ __forceinline__ void split_brick(const LsGrid& G, long long b, int& bi, int& bj,
                                            int& bk) {
  split_node(b, G.fby, G.fbz, bi, bj, bk);
}

// node of lane t in brick b, with its (i, j, k)
__device__ __forceinline__ long long brick_node(const LsGrid& G, long long b, int t, bool& valid,
                                                int& i, int& j, int& k) {
  int bi, bj, bk;
  split_brick(G, b, bi, bj, bk);
  i = bi * kBX + t / (kBY * kBZ);
  j = bj * kBY + (t / kBZ) % kBY;
  k = bk * kBZ + t % kBZ;
  valid = i < G.nx && j < G.ny && k < G.nz;
  return ((long long)i * G.ny + j) * G.nz + k;
}

// any of the flags of brick b and of its face neighbours in the flag plane f
__device__ __forceinline__ bool brick_or_nbr(const LsGrid& G, const uint8_t* __restrict__ f,
                                             long long b) {
  int bi, bj, bk;
  split_brick(G, b, bi, bj, bk);
  const long long sx = (long long)G.bny * G.bnz, sy = G.bnz;
  return f[b] || (bk > 0 && f[b - 1]) || (bk < G.bnz - 1 && f[b + 1]) ||
         (bj > 0 && f[b - sy]) || (bj < G.bny - 1 && f[b + sy]) || (bi > 0 && f[b - sx]) ||
         (bi < G.bnx - 1 && f[b + sx]);
}

__global__ void k_zero_bytes(uint8_t* __restrict__ p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Appends (f, b | tag) to the list with one counter atomic per wavefront.
__device__ __forceinline__ void list_append(bool want, int f, int bt, int2* __restrict__ list,
                                            int* __restrict__ cnt) {
  const unsigned long long m = __ballot(want);
  if (!m) return;
  const int lane = (int)(threadIdx.x & 63);
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(cnt, __popcll(m));
  base = __shfl(base, leader);
  if (want) list[base + __popcll(m & ((1ull << lane) - 1ull))] = make_int2(f, bt);
}

// Batched advection. Before the substeps, a brick is marked when one of its nodes has a nonzero
// speed or is a mask node not at its pinned value; no other node changes in any substep
// (phi - dt 0 |grad phi| = phi for a finite level set, and the pin is a no-op). The substeps
// update the marked bricks, copy their unmarked face neighbours from the source buffer to the
// destination (so every value a marked brick's stencils read is current in both ping-pong
// buffers) and leave the rest, whose phi is already final.
__global__ void k_advect_mark(const LsGrid* __restrict__ grids, int n_grids) {
  const LsGrid& G = grids[grid_of_nodes(grids, n_grids)];
  const long long N = (long long)G.nx * G.ny * G.nz;
  for (int r = 0; r < G.rounds; ++r) {
    const long long n = ((blockIdx.x - G.nblk0) * G.rounds + r) * (long long)kBS + threadIdx.x;
    if (n >= N) return;
    if (G.F[n] != 0.f || (G.mask[n] && !(G.phi[n] == G.mask_phi[n])))
      G.bfl[brick_of_node(G, n)] = 1;
  }
}

__global__ void k_advect_list(const LsGrid* __restrict__ grids, int n_grids, long long total,
                              int2* __restrict__ list, int* __restrict__ cnt) {
  const long long gb = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  bool want = false;
  int f = 0, bt = 0;
  if (gb < total) {
    f = grid_of_brick(grids, n_grids, gb);
    const LsGrid& G = grids[f];
    const long long b = gb - G.bblk0;
    if (G.count > 0) {
      if (G.bfl[b]) {
        want = true;
        bt = (int)b;
      } else if (brick_or_nbr(G, G.bfl, b)) {
        want = true;
        bt = (int)b | kHalo;
      }
    }
  }
  list_append(want, f, bt, list, cnt);
}

// substep s of every grid with more than s substeps: phi -> work on even s, work -> phi on odd
__global__ void k_advect_bricks(const LsGrid* __restrict__ grids, const int2* __restrict__ list,
                                const int* __restrict__ cnt, int s) {
  const int m = *cnt;
  for (int e = blockIdx.x; e < m; e += gridDim.x) {
    const int2 q = list[e];
    const LsGrid& G = grids[q.x];
    if (s >= G.count) continue;
    bool valid;
    int i, j, k;
    const long long n = brick_node(G, q.y & (kHalo - 1), threadIdx.x, valid, i, j, k);
    if (!valid) continue;
    const float* src = (s & 1) ? G.work : G.phi;
    float* dst = (s & 1) ? G.phi : G.work;
    if (q.y & kHalo)
      dst[n] = src[n];
    else
      advect_node(src, G.F, G.mask, G.mask_phi, dst, G.nx, G.ny, G.nz, G.inv_dx, G.dt, n, i, j,
                  k);
  }
}

// a grid with an odd substep count ends in work: copy its listed bricks back to phi
__global__ void k_advect_bricks_copyback(const LsGrid* __restrict__ grids,
                                         const int2* __restrict__ list,
                                         const int* __restrict__ cnt) {
  const int m = *cnt;
  for (int e = blockIdx.x; e < m; e += gridDim.x) {
    const int2 q = list[e];
    const LsGrid& G = grids[q.x];
    if (!(G.count & 1)) continue;
    bool valid;
    int i, j, k;
    const long long n = brick_node(G, q.y & (kHalo - 1), threadIdx.x, valid, i, j, k);
    if (valid) G.phi[n] = G.work[n];
  }
}

// Batched reinit. Flag plane k (k = 0..count) of a grid marks the bricks whose values changed in
// sweep k (plane 0: the bricks holding interface seeds). Both ping-pong buffers start at the
// initial state. A Jacobi update of a node none of whose neighbours changed since the previous
// sweep returns its current value, so sweep k lists only the bricks that changed in sweep k - 1
// or touch one that did; a brick left out keeps the previous sweep's values, which its
// destination buffer already holds (by induction: it did not change in sweep k - 1 either). The
// result equals the full sweeps' to the bit.
__global__ void k_eik_init_grids(const LsGrid* __restrict__ grids, int n_grids) {
  const LsGrid& G = grids[grid_of_nodes(grids, n_grids)];
  const long long N = (long long)G.nx * G.ny * G.nz;
  for (int r = 0; r < G.rounds; ++r) {
    const long long n = ((blockIdx.x - G.nblk0) * G.rounds + r) * (long long)kBS + threadIdx.x;
    if (n >= N) return;
    int i, j, k;
    split_node(n, G.fy, G.fz, i, j, k);
    eik_init_node(G.phi, G.U, G.fr, G.nx, G.ny, G.nz, G.dx, G.big, n, i, j, k);
    G.U2[n] = G.U[n];
    if (G.fr[n]) G.bfl[((long long)(i / kBX) * G.bny + j / kBY) * G.bnz + k / kBZ] = 1;
  }
}

__global__ void k_eik_list(const LsGrid* __restrict__ grids, int n_grids, long long total, int it,
                           int2* __restrict__ list, int* __restrict__ cnt) {
  const long long gb = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  bool want = false;
  int f = 0, bt = 0;
  if (gb < total) {
    f = grid_of_brick(grids, n_grids, gb);
    const LsGrid& G = grids[f];
    const long long b = gb - G.bblk0;
    bt = (int)b;
    want = it < G.count && brick_or_nbr(G, G.bfl + (long long)it * nbricks(G), b);
  }
  list_append(want, f, bt, list, cnt);
}

// sweep it (k = it + 1) over the listed bricks: U -> U2 on even it, U2 -> U on odd
__global__ void k_eik_jacobi_bricks(const LsGrid* __restrict__ grids,
                                    const int2* __restrict__ list, const int* __restrict__ cnt,
                                    int it) {
  const int m = *cnt;
  for (int e = blockIdx.x; e < m; e += gridDim.x) {
    const int2 q = list[e];
    const LsGrid& G = grids[q.x];
    bool valid;
    int i, j, k;
    const long long n = brick_node(G, q.y, threadIdx.x, valid, i, j, k);
    int ch = 0;
    if (valid) {
      const float* src = (it & 1) ? G.U2 : G.U;
      float* dst = (it & 1) ? G.U : G.U2;
      const float u = src[n];
      eik_jacobi_node(src, G.fr, dst, G.nx, G.ny, G.nz, G.dx, G.big, n, i, j, k);
      ch = dst[n] < u;
    }
    ch = __syncthreads_or(ch);
    if (threadIdx.x == 0 && ch) G.bfl[(long long)(it + 1) * nbricks(G) + q.y] = 1;
  }
}

__global__ void k_eik_final_grids(const LsGrid* __restrict__ grids, int n_grids) {
  const LsGrid& G = grids[grid_of_nodes(grids, n_grids)];
  const long long N = (long long)G.nx * G.ny * G.nz;
  for (int r = 0; r < G.rounds; ++r) {
    const long long n = ((blockIdx.x - G.nblk0) * G.rounds + r) * (long long)kBS + threadIdx.x;
    if (n >= N) return;
    eik_final_node(G.phi, (G.count & 1) ? G.U2 : G.U, G.clampv, n);
  }
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
  PETCH_ARG(ctx, N <= 0xffffffffLL, "advect: %lld nodes exceed 32-bit node indices", N);
  float inv = (float)(1.0 / dx);
  float fdt = (float)dt;
  float* a = phi;
  float* b = work;
  for (int s = 0; s < nsub; ++s) {
    hipLaunchKernelGGL(k_advect, dim3(petch_blocks(N, kBS)), dim3(kBS), 0, ctx->stream, a, F,
                       mask, mask_phi, b, nx, ny, nz, inv, fdt, petch::make_fastdiv((uint32_t)ny), petch::make_fastdiv((uint32_t)nz));
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
  PETCH_ARG(ctx, N <= 0xffffffffLL, "advect_step: %lld nodes exceed 32-bit node indices", N);
  hipLaunchKernelGGL(k_advect, dim3(petch_blocks(N, kBS)), dim3(kBS), 0, ctx->stream, phi, F,
                     mask, mask_phi, out, nx, ny, nz, (float)(1.0 / dx), (float)dt, petch::make_fastdiv((uint32_t)ny), petch::make_fastdiv((uint32_t)nz));
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
  if (N > 0xffffffffLL)
    return petch_set_error(ctx, PETCH_E_ARG, "reinit: %lld nodes exceed 32-bit node indices", N);
  const petch::FastDiv fy = petch::make_fastdiv((uint32_t)ny), fz = petch::make_fastdiv((uint32_t)nz);
  hipLaunchKernelGGL(k_eik_init, g, b, 0, ctx->stream, phi, U, fr, nx, ny, nz, dx, (float)big,
                     fy, fz);
  PETCH_LAUNCH_CHECK(ctx);
  for (int it = 0; it < n_iter; ++it) {
    hipLaunchKernelGGL(k_eik_jacobi, g, b, 0, ctx->stream, U, fr, U2, nx, ny, nz, dx, (float)big,
                       fy, fz);
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

extern "C" int petch_reinit_pde(petch_ctx* ctx, float* phi, int nx, int ny, int nz, double dx,
                                int n_iter) {
  PETCH_ARG(ctx, ctx && phi, "reinit_pde: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "reinit_pde: bad shape %d,%d,%d", nx, ny, nz);
  PETCH_ARG(ctx, dx > 0.0 && n_iter >= 0, "reinit_pde: dx <= 0 or n_iter < 0");
  const long long N = (long long)nx * ny * nz;
  PETCH_ARG(ctx, N <= 0xffffffffLL, "reinit_pde: %lld nodes exceed 32-bit node indices", N);
  if (n_iter == 0) return PETCH_OK;
  float* p0 = (float*)petch_scratch(ctx, SCR_LS_A, (size_t)N * sizeof(float));
  float* b = (float*)petch_scratch(ctx, SCR_LS_B, (size_t)N * sizeof(float));
  if (!p0 || !b) return PETCH_E_HIP;
  PETCH_HIP(ctx, hipMemcpyAsync(p0, phi, (size_t)N * sizeof(float), hipMemcpyDeviceToDevice,
                                ctx->stream));
  const float dx32 = (float)dx;
  const float inv = 1.0f / dx32;
  const float dtau = 0.3f * dx32;
  const petch::FastDiv fy = petch::make_fastdiv((uint32_t)ny), fz = petch::make_fastdiv((uint32_t)nz);
  const dim3 g(petch_blocks(N, kBS)), bs(kBS);
  float* a = phi;
  float* c = b;
  for (int it = 0; it < n_iter; ++it) {
    hipLaunchKernelGGL(k_reinit_pde, g, bs, 0, ctx->stream, a, p0, c, nx, ny, nz, inv, dtau, fy,
                       fz);
    PETCH_LAUNCH_CHECK(ctx);
    float* t = a; a = c; c = t;
  }
  if (a != phi)
    PETCH_HIP(ctx, hipMemcpyAsync(phi, a, (size_t)N * sizeof(float), hipMemcpyDeviceToDevice,
                                  ctx->stream));
  return PETCH_OK;
}

extern "C" int petch_gaussian_smooth(petch_ctx* ctx, float* phi, int nx, int ny, int nz,
                                     const double* weights, int radius, const uint8_t* mask,
                                     const float* mask_phi) {
  PETCH_ARG(ctx, ctx && phi && weights, "gaussian_smooth: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "gaussian_smooth: bad shape %d,%d,%d", nx, ny, nz);
  PETCH_ARG(ctx, radius >= 0 && radius <= 4096, "gaussian_smooth: radius %d out of range", radius);
  PETCH_ARG(ctx, !mask || mask_phi, "gaussian_smooth: a mask needs its pinned values");
  const long long N = (long long)nx * ny * nz;
  PETCH_ARG(ctx, N <= 0xffffffffLL, "gaussian_smooth: %lld nodes exceed 32-bit node indices", N);
  float* a = (float*)petch_scratch(ctx, SCR_LS_A, (size_t)N * sizeof(float));
  float* b = (float*)petch_scratch(ctx, SCR_LS_B, (size_t)N * sizeof(float));
  if (!a || !b) return PETCH_E_HIP;
  const double* w = (const double*)petch_upload_table(ctx, weights,
                                                      (size_t)(2 * radius + 1) * sizeof(double));
  if (!w) return PETCH_E_HIP;
  const petch::FastDiv fy = petch::make_fastdiv((uint32_t)ny), fz = petch::make_fastdiv((uint32_t)nz);
  const dim3 g(petch_blocks(N, kBS)), bs(kBS);
  hipLaunchKernelGGL(k_gauss_axis, g, bs, 0, ctx->stream, phi, a, nx, ny, nz, 0, w, radius, fy,
                     fz, (const uint8_t*)nullptr, (const float*)nullptr);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_gauss_axis, g, bs, 0, ctx->stream, a, b, nx, ny, nz, 1, w, radius, fy, fz,
                     (const uint8_t*)nullptr, (const float*)nullptr);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_gauss_axis, g, bs, 0, ctx->stream, b, phi, nx, ny, nz, 2, w, radius, fy,
                     fz, mask, mask_phi);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

namespace {

// Fills the per-grid blk0 prefix and returns the launch's workgroup count (0 on overflow).
// Sets each grid's brick counts, first workgroup of a full pass and first brick; returns the
// total brick count, or 0 if a full pass or a brick list does not fit one launch.
long long assign_grids(LsGrid* g, int n, long long* node_blocks) {
  long long b = 0, nb = 0, nodes = 0;
  for (int f = 0; f < n; ++f) nodes += (long long)g[f].nx * g[f].ny * g[f].nz;
  long long rounds = nodes / ((long long)kBS * 16384);
  rounds = rounds < 1 ? 1 : (rounds > 8 ? 8 : rounds);
  for (int f = 0; f < n; ++f) {
    g[f].bnx = (g[f].nx + kBX - 1) / kBX;
    g[f].bny = (g[f].ny + kBY - 1) / kBY;
    g[f].bnz = (g[f].nz + kBZ - 1) / kBZ;
    g[f].fy = petch::make_fastdiv((uint32_t)g[f].ny);
    g[f].fz = petch::make_fastdiv((uint32_t)g[f].nz);
    g[f].fby = petch::make_fastdiv((uint32_t)g[f].bny);
    g[f].fbz = petch::make_fastdiv((uint32_t)g[f].bnz);
    if ((long long)g[f].nx * g[f].ny * g[f].nz > 0xffffffffLL) return 0;
    g[f].nblk0 = nb;
    g[f].rounds = (int)rounds;
    nb += (long long)petch_blocks((long long)g[f].nx * g[f].ny * g[f].nz, kBS * (int)rounds);
    g[f].bblk0 = b;
    b += (long long)g[f].bnx * g[f].bny * g[f].bnz;
  }
  *node_blocks = nb;
  return (nb <= 0x7fffffffLL && b < kHalo) ? b : 0;
}

}  // namespace

extern "C" int petch_advect_batch(petch_ctx* ctx, int n, const petch_advect_item* items) {
  PETCH_ARG(ctx, ctx && (n == 0 || items), "advect_batch: null pointer");
  PETCH_ARG(ctx, n >= 0, "advect_batch: n < 0");
  if (n == 0) return PETCH_OK;
  LsGrid* g = (LsGrid*)calloc((size_t)n, sizeof(LsGrid));
  if (!g) return petch_set_error(ctx, PETCH_E_HIP, "advect_batch: out of host memory");
  int max_sub = 0;
  for (int f = 0; f < n; ++f) {
    const petch_advect_item& it = items[f];
    if (!(it.phi && it.work && it.F && it.mask && it.mask_phi) || it.nx <= 0 || it.ny <= 0 ||
        it.nz <= 0 || it.nsub < 0 || !(it.dx > 0)) {
      free(g);
      return petch_set_error(ctx, PETCH_E_ARG, "advect_batch: item %d is invalid", f);
    }
    LsGrid q = {};
    q.phi = it.phi;
    q.work = it.work;
    q.F = it.F;
    q.mask = it.mask;
    q.mask_phi = it.mask_phi;
    q.nx = it.nx;
    q.ny = it.ny;
    q.nz = it.nz;
    q.count = it.nsub;
    q.dx = it.dx;
    q.inv_dx = (float)(1.0 / it.dx);
    q.dt = (float)it.dt;
    g[f] = q;
    if (it.nsub > max_sub) max_sub = it.nsub;
  }
  long long nodeblocks = 0;
  const long long bricks = assign_grids(g, n, &nodeblocks);
  if (!bricks) {
    free(g);
    return petch_set_error(ctx, PETCH_E_ARG, "advect_batch: too many nodes for one launch");
  }
  // one flag per brick, then the list counter (int, 4-byte aligned)
  const long long fbytes = (bricks + 3) & ~3LL;
  uint8_t* bfl = (uint8_t*)petch_scratch(ctx, SCR_LS_BRICK, (size_t)fbytes + sizeof(int));
  int2* list = (int2*)petch_scratch(ctx, SCR_LS_LIST, (size_t)bricks * sizeof(int2));
  if (!bfl || !list) {
    free(g);
    return PETCH_E_HIP;
  }
  int* cnt = (int*)(bfl + fbytes);
  for (int f = 0; f < n; ++f) g[f].bfl = bfl + g[f].bblk0;
  const LsGrid* dg = (const LsGrid*)petch_upload_table(ctx, g, (size_t)n * sizeof(LsGrid));
  free(g);
  if (!dg) return PETCH_E_HIP;
  const unsigned zb = (unsigned)std::min<long long>(petch_blocks(fbytes + 4, kBS), 1024);
  hipLaunchKernelGGL(k_zero_bytes, dim3(zb), dim3(kBS), 0, ctx->stream, bfl, fbytes + 4);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_advect_mark, dim3((unsigned)nodeblocks), dim3(kBS), 0, ctx->stream, dg, n);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_advect_list, dim3(petch_blocks(bricks, kBS)), dim3(kBS), 0, ctx->stream,
                     dg, n, bricks, list, cnt);
  PETCH_LAUNCH_CHECK(ctx);
  const unsigned lw = (unsigned)std::min<long long>(bricks, kListWG);
  for (int sub = 0; sub < max_sub; ++sub) {
    hipLaunchKernelGGL(k_advect_bricks, dim3(lw), dim3(kBS), 0, ctx->stream, dg, list, cnt, sub);
    PETCH_LAUNCH_CHECK(ctx);
  }
  hipLaunchKernelGGL(k_advect_bricks_copyback, dim3(lw), dim3(kBS), 0, ctx->stream, dg, list,
                     cnt);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_reinit_fsm_batch(petch_ctx* ctx, int n, const petch_reinit_item* items) {
  PETCH_ARG(ctx, ctx && (n == 0 || items), "reinit_fsm_batch: null pointer");
  PETCH_ARG(ctx, n >= 0, "reinit_fsm_batch: n < 0");
  if (n == 0) return PETCH_OK;
  long long total = 0;
  for (int f = 0; f < n; ++f) {
    const petch_reinit_item& it = items[f];
    PETCH_ARG(ctx, it.phi && it.nx > 0 && it.ny > 0 && it.nz > 0 && it.dx > 0 && it.band > 0,
              "reinit_fsm_batch: item %d is invalid", f);
    total += (long long)it.nx * it.ny * it.nz;
  }
  float* U = (float*)petch_scratch(ctx, SCR_LS_A, (size_t)total * sizeof(float));
  float* U2 = (float*)petch_scratch(ctx, SCR_LS_B, (size_t)total * sizeof(float));
  uint8_t* fr = (uint8_t*)petch_scratch(ctx, SCR_LS_FROZEN, (size_t)total);
  if (!U || !U2 || !fr) return PETCH_E_HIP;
  LsGrid* g = (LsGrid*)calloc((size_t)n, sizeof(LsGrid));
  if (!g) return petch_set_error(ctx, PETCH_E_HIP, "reinit_fsm_batch: out of host memory");
  long long off = 0;
  int max_it = 0;
  for (int f = 0; f < n; ++f) {
    const petch_reinit_item& it = items[f];
    const int n_iter = it.n_iter < 0 ? (int)ceil(it.band / it.dx) + 4 : it.n_iter;
    LsGrid q = {};
    q.phi = it.phi;
    q.U = U + off;
    q.U2 = U2 + off;
    q.fr = fr + off;
    q.nx = it.nx;
    q.ny = it.ny;
    q.nz = it.nz;
    q.count = n_iter;
    q.dx = it.dx;
    q.big = (float)(it.band + 4.0 * it.dx);
    q.clampv = (float)(it.band + it.dx);
    g[f] = q;
    off += (long long)it.nx * it.ny * it.nz;
    if (n_iter > max_it) max_it = n_iter;
  }
  long long nodeblocks = 0;
  const long long bricks = assign_grids(g, n, &nodeblocks);
  if (!bricks) {
    free(g);
    return petch_set_error(ctx, PETCH_E_ARG, "reinit_fsm_batch: too many nodes for one launch");
  }
  // flag planes 0..count of every grid (grid f's after those of the grids before it), then one
  // list counter per sweep
  long long nflags = 0;
  for (int f = 0; f < n; ++f)
    nflags += (long long)(g[f].count + 1) * g[f].bnx * g[f].bny * g[f].bnz;
  const long long fbytes = (nflags + 3) & ~3LL;
  const long long zbytes = fbytes + (long long)(max_it + 1) * (long long)sizeof(int);
  uint8_t* bfl = (uint8_t*)petch_scratch(ctx, SCR_LS_BRICK, (size_t)zbytes);
  int2* list = (int2*)petch_scratch(ctx, SCR_LS_LIST, (size_t)bricks * sizeof(int2));
  if (!bfl || !list) {
    free(g);
    return PETCH_E_HIP;
  }
  int* cnt = (int*)(bfl + fbytes);
  long long foff = 0;
  for (int f = 0; f < n; ++f) {
    g[f].bfl = bfl + foff;
    foff += (long long)(g[f].count + 1) * g[f].bnx * g[f].bny * g[f].bnz;
  }
  const LsGrid* dg = (const LsGrid*)petch_upload_table(ctx, g, (size_t)n * sizeof(LsGrid));
  free(g);
  if (!dg) return PETCH_E_HIP;
  const unsigned zb = (unsigned)std::min<long long>(petch_blocks(zbytes, kBS), 1024);
  hipLaunchKernelGGL(k_zero_bytes, dim3(zb), dim3(kBS), 0, ctx->stream, bfl, zbytes);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_eik_init_grids, dim3((unsigned)nodeblocks), dim3(kBS), 0, ctx->stream, dg,
                     n);
  PETCH_LAUNCH_CHECK(ctx);
  const unsigned lw = (unsigned)std::min<long long>(bricks, kListWG);
  for (int it = 0; it < max_it; ++it) {
    hipLaunchKernelGGL(k_eik_list, dim3(petch_blocks(bricks, kBS)), dim3(kBS), 0, ctx->stream, dg,
                       n, bricks, it, list, cnt + it);
    PETCH_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_eik_jacobi_bricks, dim3(lw), dim3(kBS), 0, ctx->stream, dg, list,
                       cnt + it, it);
    PETCH_LAUNCH_CHECK(ctx);
  }
  hipLaunchKernelGGL(k_eik_final_grids, dim3((unsigned)nodeblocks), dim3(kBS), 0, ctx->stream, dg,
                     n);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
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

// ------------------------------------------------------------------------------ reinit adjoint
// The adjoint of the FSM reinit as the piecewise-smooth map its launches compute: every discrete
// choice (interface flags, the smaller neighbour per axis, the active Godunov terms, whether a
// sweep lowers a node, the clamp at band + dx) is the forward's, recomputed from the sweep
// states, which the adjoint rebuilds with the forward's own kernels. A node that sweep k lowers
// took eik_solve of its three axis minima at state k; the root x of sum_d (x - a_d)^2 = dx^2 over
// the active terms moves with a_d by (x - a_d) / sum_e (x - a_e). The frozen seeds
// |phi| / max(|grad phi|_central, 0.5) pass their derivative to the node and its six neighbours.

// Godunov root of the three axis minima in float64 and its weight on each axis (0 off the
// active set); the same staging as eik_solve, so the active set is the forward's.
__device__ __forceinline__ void eik_weights(float f1, float f2, float f3, double dx, double* w) {
  double a[3] = {f1, f2, f3};
  double s[3] = {f1, f2, f3}, t;
  if (s[0] > s[1]) { t = s[0]; s[0] = s[1]; s[1] = t; }
  if (s[1] > s[2]) { t = s[1]; s[1] = s[2]; s[2] = t; }
  if (s[0] > s[1]) { t = s[0]; s[0] = s[1]; s[1] = t; }
  double x = s[0] + dx;
  if (x > s[1]) {
    double b2 = s[1] - s[0], b3 = s[2] - s[0];
    double y = 0.5 * (b2 + sqrt(fmax(2.0 * dx * dx - b2 * b2, 0.0)));
    if (s[0] + y <= s[2]) {
      x = s[0] + y;
    } else {
      double s1 = b2 + b3;
      x = s[0] + (s1 + sqrt(fmax(s1 * s1 - 3.0 * (b2 * b2 + b3 * b3 - dx * dx), 0.0))) / 3.0;
    }
  }
  double den = 0.0;
  for (int d = 0; d < 3; ++d) {
    w[d] = a[d] < x ? x - a[d] : 0.0;
    den += w[d];
  }
  for (int d = 0; d < 3; ++d) w[d] = den > 0.0 ? w[d] / den : 0.0;
}

// Per node of sweep k's state U: whether the sweep lowered it, and the weights it puts on the
// smaller neighbour of each axis (W.w = bit 0 took, bits 1..3 "the upper neighbour" per axis).
__global__ void k_eik_adj_weights(const float* __restrict__ U, const uint8_t* __restrict__ frozen,
                                  float4* __restrict__ W, int nx, int ny, int nz, double dx,
                                  float big) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long N = (long long)nx * ny * nz;
  if (n >= N) return;
  float4 out = make_float4(0.f, 0.f, 0.f, __int_as_float(0));
  if (!frozen[n]) {
    int k = (int)(n % nz);
    int j = (int)((n / nz) % ny);
    int i = (int)(n / ((long long)nz * ny));
    long long sx = (long long)ny * nz, sy = nz;
    float lo[3] = {i >= 1 ? U[n - sx] : big, j >= 1 ? U[n - sy] : big, k >= 1 ? U[n - 1] : big};
    float hi[3] = {i <= nx - 2 ? U[n + sx] : big, j <= ny - 2 ? U[n + sy] : big,
                   k <= nz - 2 ? U[n + 1] : big};
    float a[3];
    int bits = 0;
    for (int d = 0; d < 3; ++d) {
      a[d] = fminf(lo[d], hi[d]);
      if (!(lo[d] <= hi[d])) bits |= 2 << d;     // the lower side on a tie
    }
    float cand = eik_solve(a[0], a[1], a[2], dx);
    if (cand < U[n]) {
      double w[3];
      eik_weights(a[0], a[1], a[2], dx, w);
      out = make_float4((float)w[0], (float)w[1], (float)w[2], __int_as_float(bits | 1));
    }
  }
  W[n] = out;
}

// G_k = transpose of sweep k applied to G_{k+1}: a node keeps its own adjoint unless the sweep
// lowered it, and gathers w * G from every neighbour that took it as its axis minimum.
__global__ void k_eik_adj_sweep(const float4* __restrict__ W, const float* __restrict__ Gin,
                                float* __restrict__ Gout, int nx, int ny, int nz) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long N = (long long)nx * ny * nz;
  if (n >= N) return;
  int k = (int)(n % nz);
  int j = (int)((n / nz) % ny);
  int i = (int)(n / ((long long)nz * ny));
  const long long st[3] = {(long long)ny * nz, nz, 1};
  const int idx[3] = {i, j, k}, ext[3] = {nx, ny, nz};
  float4 me = W[n];
  float g = (__float_as_int(me.w) & 1) ? 0.f : Gin[n];
  for (int d = 0; d < 3; ++d) {
    if (idx[d] + 1 < ext[d]) {               // the upper neighbour took n if its choice is "lower"
      const long long m = n + st[d];
      float4 q = W[m];
      int b = __float_as_int(q.w);
      float wd = d == 0 ? q.x : (d == 1 ? q.y : q.z);
      if ((b & 1) && !(b & (2 << d))) g += wd * Gin[m];
    }
    if (idx[d] >= 1) {                        // the lower neighbour took n if its choice is "upper"
      const long long m = n - st[d];
      float4 q = W[m];
      int b = __float_as_int(q.w);
      float wd = d == 0 ? q.x : (d == 1 ? q.y : q.z);
      if ((b & 1) && (b & (2 << d))) g += wd * Gin[m];
    }
  }
  Gout[n] = g;
}

// The clamp and the sign of the output: G = gout * sign(phi0) where U < band + dx, else 0.
__global__ void k_eik_adj_final(const float* __restrict__ phi0, const float* __restrict__ U,
                                const float* __restrict__ gout, float* __restrict__ G,
                                long long N, float clampv) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float p = phi0[n];
  float s = (p < 0.f) ? -1.f : 1.f;
  G[n] = (p == p && U[n] < clampv) ? gout[n] * s : 0.f;
}

// The seed terms of frozen node m at G0 = G[m]: d D / d phi0[m] and the coefficient c_d of the
// central difference on axis d (d D / d phi0[m +- e_d] = +- c_d).
__device__ __forceinline__ void eik_seed_terms(const float* __restrict__ phi0, long long m, int i,
                                               int j, int k, int nx, int ny, int nz, double dx,
                                               float G0, double* self, double* cd) {
  long long sx = (long long)ny * nz, sy = nz;
  double c = phi0[m];
  double oxm = phi0[i > 0 ? m - sx : m], oxp = phi0[i < nx - 1 ? m + sx : m];
  double oym = phi0[j > 0 ? m - sy : m], oyp = phi0[j < ny - 1 ? m + sy : m];
  double ozm = phi0[k > 0 ? m - 1 : m], ozp = phi0[k < nz - 1 ? m + 1 : m];
  double h = 0.5 / dx;
  double g[3] = {(oxp - oxm) * h, (oyp - oym) * h, (ozp - ozm) * h};
  double gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  double grad0 = fmax(gn, 0.5);
  *self = 0.0;
  cd[0] = cd[1] = cd[2] = 0.0;
  if (!(fabs(c) / grad0 < 1.8 * dx)) return;          // capped seed: constant
  double sg = c > 0.0 ? 1.0 : (c < 0.0 ? -1.0 : 0.0);
  *self = G0 * sg / grad0;
  if (gn > 0.5) {
    double coef = -G0 * fabs(c) / (grad0 * grad0 * gn);
    for (int d = 0; d < 3; ++d) cd[d] = coef * g[d] * h;
  }
}

// d loss / d phi0 from G0 (the adjoint of the initial state): only frozen nodes depend on phi0.
__global__ void k_eik_adj_init(const float* __restrict__ phi0, const uint8_t* __restrict__ frozen,
                               const float* __restrict__ G, float* __restrict__ gphi, int nx,
                               int ny, int nz, double dx) {
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long N = (long long)nx * ny * nz;
  if (n >= N) return;
  int k = (int)(n % nz);
  int j = (int)((n / nz) % ny);
  int i = (int)(n / ((long long)nz * ny));
  const long long st[3] = {(long long)ny * nz, nz, 1};
  const int ext[3] = {nx, ny, nz};
  int idx[3] = {i, j, k};
  double acc = 0.0, self, cd[3];
  if (frozen[n]) {
    eik_seed_terms(phi0, n, i, j, k, nx, ny, nz, dx, G[n], &self, cd);
    acc += self;
    for (int d = 0; d < 3; ++d) {            // a clamped neighbour at the edge is n itself
      if (idx[d] == ext[d] - 1) acc += cd[d];
      if (idx[d] == 0) acc -= cd[d];
    }
  }
  for (int d = 0; d < 3; ++d) {
    for (int side = -1; side <= 1; side += 2) {
      int q = idx[d] + side;
      if (q < 0 || q >= ext[d]) continue;
      const long long m = n + side * st[d];
      if (!frozen[m]) continue;
      int mi[3] = {i, j, k};
      mi[d] = q;
      eik_seed_terms(phi0, m, mi[0], mi[1], mi[2], nx, ny, nz, dx, G[m], &self, cd);
      // n is m's +1 neighbour when m sits below it (side = -1), its -1 neighbour otherwise
      acc += side < 0 ? cd[d] : -cd[d];
    }
  }
  gphi[n] = (float)acc;
}

extern "C" int petch_reinit_fsm_adjoint(petch_ctx* ctx, const float* phi0, int nx, int ny, int nz,
                                        double dx, double band, int n_iter, const float* gout,
                                        float* gphi) {
  PETCH_ARG(ctx, ctx && phi0 && gout && gphi, "reinit_fsm_adjoint: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "reinit_fsm_adjoint: bad shape %d,%d,%d", nx, ny,
            nz);
  PETCH_ARG(ctx, dx > 0 && band > 0, "reinit_fsm_adjoint: dx and band must be > 0");
  PETCH_ARG(ctx, gphi != gout && gphi != phi0, "reinit_fsm_adjoint: gphi must not alias");
  const long long N = (long long)nx * ny * nz;
  const double big = band + 4.0 * dx;
  if (n_iter < 0) n_iter = (int)ceil(band / dx) + 4;
  const size_t slab = round256((size_t)N * sizeof(float));
  char* S = (char*)petch_scratch(ctx, SCR_EIK_STATES, slab * (size_t)(n_iter + 1));
  float4* W = (float4*)petch_scratch(ctx, SCR_EIK_W, (size_t)N * sizeof(float4));
  char* G = (char*)petch_scratch(ctx, SCR_EIK_G, 2 * slab);
  uint8_t* fr = (uint8_t*)petch_scratch(ctx, SCR_LS_FROZEN, N);
  if (!S || !W || !G || !fr) return PETCH_E_HIP;
  auto state = [&](int k) { return (float*)(S + slab * (size_t)k); };
  float* Ga = (float*)G;
  float* Gb = (float*)(G + slab);
  dim3 g(petch_blocks(N, kBS)), b(kBS);
  // the forward's states, with the forward's kernels
  PETCH_ARG(ctx, N <= 0xffffffffLL, "reinit_fsm_adjoint: %lld nodes exceed 32-bit node indices", N);
  const petch::FastDiv fy = petch::make_fastdiv((uint32_t)ny), fz = petch::make_fastdiv((uint32_t)nz);
  hipLaunchKernelGGL(k_eik_init, g, b, 0, ctx->stream, phi0, state(0), fr, nx, ny, nz, dx,
                     (float)big, fy, fz);
  PETCH_LAUNCH_CHECK(ctx);
  for (int it = 0; it < n_iter; ++it) {
    hipLaunchKernelGGL(k_eik_jacobi, g, b, 0, ctx->stream, state(it), fr, state(it + 1), nx, ny,
                       nz, dx, (float)big, fy, fz);
    PETCH_LAUNCH_CHECK(ctx);
  }
  hipLaunchKernelGGL(k_eik_adj_final, g, b, 0, ctx->stream, phi0, state(n_iter), gout, Ga, N,
                     (float)(band + dx));
  PETCH_LAUNCH_CHECK(ctx);
  for (int it = n_iter - 1; it >= 0; --it) {
    hipLaunchKernelGGL(k_eik_adj_weights, g, b, 0, ctx->stream, state(it), fr, W, nx, ny, nz, dx,
                       (float)big);
    PETCH_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_eik_adj_sweep, g, b, 0, ctx->stream, W, Ga, Gb, nx, ny, nz);
    PETCH_LAUNCH_CHECK(ctx);
    float* t = Ga; Ga = Gb; Gb = t;
  }
  hipLaunchKernelGGL(k_eik_adj_init, g, b, 0, ctx->stream, phi0, fr, Ga, gphi, nx, ny, nz, dx);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}
