// Deterministic neutral transports behind Flags.neutral_transport (the alternatives to the
// Monte-Carlo neutrals; the ions stay Monte-Carlo):
//   petch_knudsen_face_flux  1-D Knudsen conductance tail (knudsen.py:185-241): per-slice gas
//                            area and perimeter of the level set, the floor faces' sink per
//                            slice, a tridiagonal solve down the feature, each face reading its
//                            slice's concentration
//   petch_dda_neutral_flux   discrete-ordinates march (dda.py:138-220): every face gathers the
//                            sky and the walls' re-emission over fixed sphere directions by
//                            marching the grid, iterated with the per-cell re-emission field,
//                            then an Aitken step
// Everything runs on the device; the only host work is the direction set and the launch
// sequence. Both take two sticking coefficients (the F and O species of one coverage round)
// and share the per-call geometry between them.
#include "common.h"
#include "scan.h"

#include <math.h>
#include <vector>

namespace {

constexpr int kBS = 256;

// ------------------------------------------------------------------ Knudsen -----------------

// One workgroup per z-slice k (z = k dx): gas area and sidewall perimeter (knudsen.py:24-71).
// Hole: gas cells dx^2 and gas/solid transitions along x and y times dx (at least dx). Trench
// (along y): the mean open x-width of the rows with any gas; perimeter 2. Slices at or above
// sub_top, or without gas, stay 0. Also clears the slice's two sink accumulators.
__global__ void k_kn_slices(const float* __restrict__ phi, int nx, int ny, int nz, double dx,
                            double sub_top, int trench, double* __restrict__ area,
                            double* __restrict__ perim, double* __restrict__ loss) {
  const int k = blockIdx.x;
  const bool below = (double)k * dx < sub_top;
  long long a = 0, b = 0;   // hole: gas cells, transitions; trench: sum of row counts, open rows
  if (below) {
    if (trench) {
      for (int j = threadIdx.x; j < ny; j += kBS) {
        int c = 0;
        for (int i = 0; i < nx; ++i) c += phi[((long long)i * ny + j) * nz + k] < 0.f;
        if (c > 0) {
          a += c;
          b += 1;
        }
      }
    } else {
      for (long long q = threadIdx.x; q < (long long)nx * ny; q += kBS) {
        const int i = (int)(q / ny), j = (int)(q % ny);
        const bool g = phi[((long long)i * ny + j) * nz + k] < 0.f;
        a += g;
        if (i + 1 < nx) b += g != (phi[((long long)(i + 1) * ny + j) * nz + k] < 0.f);
        if (j + 1 < ny) b += g != (phi[((long long)i * ny + j + 1) * nz + k] < 0.f);
      }
    }
  }
  __shared__ long long sa[kBS], sb[kBS];
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int s = kBS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double ar = 0.0, pe = 0.0;
    if (below && sa[0] > 0) {
      if (trench) {
        ar = (double)sa[0] * dx / (double)sb[0];
        pe = 2.0;
      } else {
        ar = (double)sa[0] * dx * dx;
        pe = fmax((double)sb[0] * dx, dx);
      }
    }
    area[k] = ar;
    perim[k] = pe;
    loss[k] = 0.0;
    loss[nz + k] = 0.0;
  }
}

__device__ __forceinline__ bool kn_valid(const double* __restrict__ area,
                                         const double* __restrict__ perim, int k, double dx,
                                         double sub_top) {
  return (double)k * dx < sub_top && area[k] > 0.0 && perim[k] > 0.0;
}

struct KnFace {
  double dx, sub_top, cx, cy, band, beta[2];
  int nz, trench, footprint, sink_field;
};

// One workgroup: the mean bare coverage of the open-field faces (centroid z > sub_top - dx/2),
// or -1 without any (knudsen_sink = 'field', threed.py:1193-1197).
__global__ void k_kn_field_mean(const float* __restrict__ cen, const float* __restrict__ bare,
                                int F, double sub_top, double dx, double* __restrict__ out) {
  double s = 0.0;
  long long c = 0;
  for (int f = threadIdx.x; f < F; f += kBS) {
    if ((double)cen[3 * f + 2] > sub_top - 0.5 * dx) {
      s += (double)bare[f];
      ++c;
    }
  }
  __shared__ double ss[kBS];
  __shared__ long long sc[kBS];
  ss[threadIdx.x] = s;
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int h = kBS / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      ss[threadIdx.x] += ss[threadIdx.x + h];
      sc[threadIdx.x] += sc[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sc[0] > 0 ? ss[0] / (double)sc[0] : -1.0;
}

// Per face: its slice (rounded z / dx, clamped, snapped to the nearest valid slice with ties to
// the lower one: knudsen.py:154-163, 206-212), and, for the floor faces (into-gas normal z >
// 0.7) below sub_top inside the footprint, 0.25 clip(s) A into the slice's sink of each species.
__global__ void k_kn_snap(const float* __restrict__ cen, const float* __restrict__ gas_n,
                          const float* __restrict__ areas, const float* __restrict__ bare, int F,
                          KnFace p, const double* __restrict__ area,
                          const double* __restrict__ perim, const double* __restrict__ field_mean,
                          int* __restrict__ snap, double* __restrict__ loss) {
  const int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  const double cz = (double)cen[3 * f + 2];
  int iz = (int)fmin(fmax(rint(cz / p.dx), 0.0), (double)(p.nz - 1));
  int sn = -1;
  for (int d = 0; d < p.nz && sn < 0; ++d) {
    if (iz - d >= 0 && kn_valid(area, perim, iz - d, p.dx, p.sub_top)) sn = iz - d;
    else if (iz + d < p.nz && kn_valid(area, perim, iz + d, p.dx, p.sub_top)) sn = iz + d;
  }
  snap[f] = sn;
  if (sn < 0 || !(cz < p.sub_top) || !((double)gas_n[3 * f + 2] > 0.7)) return;
  if (p.footprint) {
    const double x = (double)cen[3 * f], y = (double)cen[3 * f + 1];
    const bool in = p.trench ? fabs(x - p.cx) <= p.band : hypot(x - p.cx, y - p.cy) <= p.band;
    if (!in) return;
  }
  double sb = (double)bare[f];
  if (p.sink_field && *field_mean >= 0.0) sb = fmin(sb, *field_mean);
  const double A = fmax((double)areas[f], 1.0e-9);
  for (int s = 0; s < 2; ++s) {
    const double st = fmin(fmax(sb * p.beta[s], 0.0), 1.0);
    atomicAdd(loss + (long long)s * p.nz + sn, 0.25 * st * A);
  }
}

struct KnSolve {
  double dx, sub_top, Ly, wls, band_W, pass_frac, ar_pass, W;
  int nz, trench, front_loss;
};

// One thread per species: the conductance chain of the valid slices from the opening down
// (C = (2/3) r a / dx, r = max(2a/p, dx); adjacent slices linked by 2 C0 C1 / (C0 + C1)), the
// unit-concentration opening, the scaled sink, and its Thomas solve with the 1e-30 pivot guard
// (knudsen.py:93-151, 74-90). prof receives the clipped concentration of every slice (1 off the
// chain). w: 4 nz doubles of workspace per species.
__global__ void k_kn_solve(const double* __restrict__ area, const double* __restrict__ perim,
                           const double* __restrict__ loss, KnSolve p, double* __restrict__ prof,
                           double* __restrict__ w) {
  const int s = threadIdx.x;
  if (s >= 2) return;
  const int nz = p.nz;
  const double* L = loss + (long long)s * nz;
  double* P = prof + (long long)s * nz;
  double* C = w + (long long)s * 4 * nz;      // conductance of chain position i
  double* diag = C + nz;
  double* off = diag + nz;                    // link between positions i and i + 1
  double* rhs = off + nz;
  for (int k = 0; k < nz; ++k) P[k] = 1.0;
  // chain positions: valid slices from the highest (the opening) down
  int m = 0;
  int prev = -1;
  for (int k = nz - 1; k >= 0; --k) {
    if (!kn_valid(area, perim, k, p.dx, p.sub_top)) continue;
    const double a = area[k], pe = perim[k];
    const double r = fmax(2.0 * a / fmax(pe, 1.0e-30), p.dx);
    C[m] = (2.0 / 3.0) * r * a / p.dx;
    double sc = p.wls;
    if (p.front_loss) {
      const double ar = fmax((p.sub_top - (double)k * p.dx) / fmax(p.W, 1.0e-9), 0.0);
      const double g = exp(-fmax(ar - p.band_W, 0.0) / fmax(p.ar_pass, 1.0e-9));
      sc = fmax(p.wls, 0.0) * (p.pass_frac + (1.0 - p.pass_frac) * g);
    }
    double sink = L[k];
    if (p.trench) sink /= fmax(p.Ly, 1.0e-30);
    diag[m] = fmax(sc, 0.0) * fmax(sink, 0.0);
    rhs[m] = 0.0;
    if (m > 0) off[m - 1] = (prev - k == 1) ? 2.0 * C[m - 1] * C[m] / fmax(C[m - 1] + C[m], 1.0e-30) : 0.0;
    prev = k;
    ++m;
  }
  if (m == 0) return;
  diag[0] += C[0];
  rhs[0] = C[0];
  for (int i = 0; i + 1 < m; ++i) {
    diag[i] += off[i];
    diag[i + 1] += off[i];
  }
  for (int i = 0; i < m; ++i) diag[i] = fmax(diag[i], 1.0e-30);
  // Thomas: lower = upper = -off
  for (int i = 1; i < m; ++i) {
    const double wv = -off[i - 1] / fmax(diag[i - 1], 1.0e-30);
    diag[i] -= wv * -off[i - 1];
    rhs[i] -= wv * rhs[i - 1];
  }
  // back substitution into C (no longer needed), then scatter to the slices
  C[m - 1] = rhs[m - 1] / fmax(diag[m - 1], 1.0e-30);
  for (int i = m - 2; i >= 0; --i) C[i] = (rhs[i] - -off[i] * C[i + 1]) / fmax(diag[i], 1.0e-30);
  int i = 0;
  for (int k = nz - 1; k >= 0; --k) {
    if (!kn_valid(area, perim, k, p.dx, p.sub_top)) continue;
    P[k] = fmin(fmax(C[i], 0.0), 1.0);
    ++i;
  }
}

__global__ void k_kn_face(const float* __restrict__ cen, const int* __restrict__ snap, int F,
                          int nz, double sub_top, const double* __restrict__ prof,
                          float* __restrict__ m0, float* __restrict__ m1) {
  const int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  const int s = snap[f];
  const bool use = s >= 0 && (double)cen[3 * f + 2] < sub_top;
  const double a = use ? prof[s] : 1.0, b = use ? prof[nz + s] : 1.0;
  m0[f] = (float)fmin(fmax(a, 0.0), 8.0);
  if (m1) m1[f] = (float)fmin(fmax(b, 0.0), 8.0);
}

// ------------------------------------------------------------------ DDA ---------------------

struct DdaGrid {
  const float* phi;
  int nx, ny, nz;
  float dx;
  float z_top, Ly;
  int max_steps, periodic_y;
};

__device__ __forceinline__ float node(const DdaGrid& g, int i, int j, int k) {
  return g.phi[((long long)i * g.ny + j) * g.nz + k];
}

__device__ __forceinline__ int round_cell(float q, float dx, int n) {
#pragma clang fp contract(off)
  const int c = (int)(q / dx + 0.5f);
  return c < 0 ? 0 : (c > n - 1 ? n - 1 : c);
}

// The value one direction ray brings back to origin o (dda.py:44-93, in the reference's float32
// order, no contraction): from o + 1.01 dx d in steps of dx d, src_top once z >= z_top; at the
// first point whose trilinear phi (coordinates clamped to [0, n - 1.001]) is > 0 the remit of
// the first solid rounded cell among six half steps on, else of the point's rounded cell; 0
// after max_steps. remit may be NULL (all zero).
__device__ float dda_ray(const DdaGrid& g, float ox, float oy, float oz, float dxv, float dyv,
                         float dzv, const float* __restrict__ remit, float src_top) {
#pragma clang fp contract(off)
  const float dx = g.dx;
  float px = ox + dxv * dx * 1.01f;
  float py = oy + dyv * dx * 1.01f;
  float pz = oz + dzv * dx * 1.01f;
  const float sx = dxv * dx, sy = dyv * dx, sz = dzv * dx;
  const float ux = (float)g.nx - 1.001f, uy = (float)g.ny - 1.001f, uz = (float)g.nz - 1.001f;
  for (int s = 0; s < g.max_steps; ++s) {
    if (pz >= g.z_top) return src_top;
    const float u = fminf(fmaxf(px / dx, 0.f), ux);
    const float v = fminf(fmaxf(py / dx, 0.f), uy);
    const float w = fminf(fmaxf(pz / dx, 0.f), uz);
    const int i0 = (int)u, j0 = (int)v, k0 = (int)w;
    const float fu = u - (float)i0, fv = v - (float)j0, fw = w - (float)k0;
    const float gu = 1.f - fu, gv = 1.f - fv, gw = 1.f - fw;
    const float p00 = node(g, i0, j0, k0) * gu + node(g, i0 + 1, j0, k0) * fu;
    const float p10 = node(g, i0, j0 + 1, k0) * gu + node(g, i0 + 1, j0 + 1, k0) * fu;
    const float p01 = node(g, i0, j0, k0 + 1) * gu + node(g, i0 + 1, j0, k0 + 1) * fu;
    const float p11 = node(g, i0, j0 + 1, k0 + 1) * gu + node(g, i0 + 1, j0 + 1, k0 + 1) * fu;
    const float pv = (p00 * gv + p10 * fv) * gw + (p01 * gv + p11 * fv) * fw;
    if (pv > 0.f) {
      if (!remit) return 0.f;
      float qx = px, qy = py, qz = pz;
      for (int e = 0; e < 6; ++e) {
        const int ix = round_cell(qx, dx, g.nx), iy = round_cell(qy, dx, g.ny),
                  iz = round_cell(qz, dx, g.nz);
        if (node(g, ix, iy, iz) > 0.f) return remit[((long long)ix * g.ny + iy) * g.nz + iz];
        qx = qx + sx * 0.5f;
        qy = qy + sy * 0.5f;
        qz = qz + sz * 0.5f;
      }
      const int ix = round_cell(px, dx, g.nx), iy = round_cell(py, dx, g.ny),
                iz = round_cell(pz, dx, g.nz);
      return remit[((long long)ix * g.ny + iy) * g.nz + iz];
    }
    px = px + sx;
    py = py + sy;
    pz = pz + sz;
    if (g.periodic_y) {
      if (py < 0.f) py = py + g.Ly;
      if (py >= g.Ly) py = py - g.Ly;
    }
  }
  return 0.f;
}

// One wavefront per face, its lanes over the directions: sum_k acc_k c_k / max(sum_k acc_k,
// 1e-12), acc_k = max(n . d_k, 0), directions with acc <= 1e-6 not marched (dda.py:36-95).
// The sums are float64 (the reference's are float32, so the two differ in the last bits).
__global__ void __launch_bounds__(kBS) k_dda_gather(DdaGrid g, const float* __restrict__ cen,
                                                    const float* __restrict__ nrm, int F,
                                                    const float* __restrict__ dirs, int nd,
                                                    const float* __restrict__ remit,
                                                    float src_top, float* __restrict__ out) {
#pragma clang fp contract(off)
  const int f = blockIdx.x * (kBS / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double num = 0.0, den = 0.0;
  if (f < F) {
    const float ox = cen[3 * f], oy = cen[3 * f + 1], oz = cen[3 * f + 2];
    const float nx = nrm[3 * f], ny = nrm[3 * f + 1], nz = nrm[3 * f + 2];
    for (int k = lane; k < nd; k += 64) {
      const float dxv = dirs[3 * k], dyv = dirs[3 * k + 1], dzv = dirs[3 * k + 2];
      const float acc = fmaxf(nx * dxv + ny * dyv + nz * dzv, 0.f);
      den += (double)acc;
      if (acc > 1.0e-6f) num += (double)acc * (double)dda_ray(g, ox, oy, oz, dxv, dyv, dzv, remit, src_top);
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    num += __shfl_xor(num, d, 64);
    den += __shfl_xor(den, d, 64);
  }
  if (f < F && lane == 0) out[f] = (float)(num / fmax(den, 1.0e-12));
}

__global__ void k_dda_zero(double* __restrict__ a, double* __restrict__ b, uint8_t* __restrict__ c,
                           long long n) {
  for (long long i = (long long)blockIdx.x * kBS + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBS) {
    if (a) a[i] = 0.0;
    if (b) b[i] = 0.0;
    if (c) c[i] = 0;
  }
}

// Per face: the solid cell 0.7 dx under it against its into-gas normal (rounded, clamped), its
// area added to the cell's, the cell marked as seeded (dda.py:157-176).
__global__ void k_dda_cells(const float* __restrict__ cen, const float* __restrict__ nrm,
                            const float* __restrict__ areas, int F, int nx, int ny, int nz,
                            double dx, int* __restrict__ cell, double* __restrict__ carea,
                            uint8_t* __restrict__ seeded) {
  const int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  const double inv = 1.0 / dx;
  int c3[3];
  const int n3[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    const double sp = (double)cen[3 * f + a] - 0.7 * dx * (double)nrm[3 * f + a];
    long long q = (long long)(sp * inv + 0.5);
    c3[a] = (int)(q < 0 ? 0 : (q > n3[a] - 1 ? n3[a] - 1 : q));
  }
  const int c = (c3[0] * ny + c3[1]) * nz + c3[2];
  cell[f] = c;
  atomicAdd(carea + c, fmax((double)areas[f], 1.0e-9));
  seeded[c] = 1;
}

// Per face: its re-emitted intensity A (1 - s) max(vf, 0) into its cell (dda.py:187).
__global__ void k_dda_scatter(const int* __restrict__ cell, const float* __restrict__ areas,
                              const float* __restrict__ bare, double beta,
                              const double* __restrict__ vf, int F, double* __restrict__ R) {
  const int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  const double s = fmin(fmax((double)bare[f] * beta, 0.0), 1.0);
  const double alpha = fmin(fmax(1.0 - s, 0.0), 1.0);
  atomicAdd(R + cell[f], fmax((double)areas[f], 1.0e-9) * alpha * fmax(vf[f], 0.0));
}

// Per cell: R = sum / max(area, 1e-12) (the area-weighted mean of the cell's faces).
__global__ void k_dda_mean(double* __restrict__ R, const double* __restrict__ carea, long long n) {
  for (long long i = (long long)blockIdx.x * kBS + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBS)
    R[i] = R[i] / fmax(carea[i], 1.0e-12);
}

// Per cell: the field the march reads (float32). A solid unseeded cell with seeded neighbours
// (the six, wrapping around the grid) takes their mean, summed in the order -x, +x, -y, +y, -z,
// +z of the reference's rolls (dda.py:196-207); every other cell keeps R.
__global__ void k_dda_fill(const float* __restrict__ phi, const double* __restrict__ R,
                           const uint8_t* __restrict__ seeded, int nx, int ny, int nz,
                           float* __restrict__ out) {
  const long long n = (long long)nx * ny * nz;
  for (long long c = (long long)blockIdx.x * kBS + threadIdx.x; c < n;
       c += (long long)gridDim.x * kBS) {
    double v = R[c];
    if (phi[c] > 0.f && !seeded[c]) {
      const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((long long)nz * ny));
      const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
      const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
      const int km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0 : k + 1;
      const long long nb[6] = {((long long)im * ny + j) * nz + k, ((long long)ip * ny + j) * nz + k,
                               ((long long)i * ny + jm) * nz + k, ((long long)i * ny + jp) * nz + k,
                               ((long long)i * ny + j) * nz + km, ((long long)i * ny + j) * nz + kp};
      double num = 0.0;
      int cnt = 0;
      for (int q = 0; q < 6; ++q) {
        if (seeded[nb[q]]) {
          num += R[nb[q]];
          ++cnt;
        }
      }
      if (cnt > 0) v = num / (double)cnt;
    }
    out[c] = (float)v;
  }
}

// vf = direct (+ the gathered re-emission g, when given)
__global__ void k_dda_update(const float* __restrict__ direct, const float* __restrict__ g,
                             int F, double* __restrict__ vf) {
  const int f = blockIdx.x * kBS + threadIdx.x;
  if (f < F) vf[f] = g ? (double)direct[f] + (double)g[f] : (double)direct[f];
}

// Aitken on the last three iterates (dda.py:212-218) when there are three, then clip to [0, 8].
__global__ void k_dda_final(const double* __restrict__ vf, const double* __restrict__ p1,
                            const double* __restrict__ p2, int F, float* __restrict__ out) {
  const int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  double v = vf[f];
  if (p2) {
    const double d1 = v - p1[f], d0 = p1[f] - p2[f];
    double r = fabs(d0) > 1.0e-12 ? d1 / d0 : 0.0;
    r = fmin(fmax(r, 0.0), 0.995);
    v = v + d1 * r / (1.0 - r);
  }
  out[f] = (float)fmin(fmax(fmax(v, 0.0), 0.0), 8.0);
}

unsigned grid_of(long long n) {
  const long long b = (n + kBS - 1) / kBS;
  return (unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

}  // namespace

extern "C" int petch_knudsen_face_flux(petch_ctx* ctx, const float* phi, int nx, int ny, int nz,
                                       const float* centroids, const float* gas_normals,
                                       const float* areas, const float* bare, int F,
                                       const petch_knudsen_params* p, float* m_a, float* m_b) {
  PETCH_ARG(ctx, ctx && phi && p && m_a, "knudsen_face_flux: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0 && F >= 0, "knudsen_face_flux: bad shape");
  PETCH_ARG(ctx, p->dx > 0.0, "knudsen_face_flux: dx <= 0");
  if (F == 0) return PETCH_OK;
  PETCH_ARG(ctx, centroids && gas_normals && areas && bare, "knudsen_face_flux: null face array");
  // slices: area, perimeter, two sink accumulators, two profiles, 8 nz of solve workspace, the
  // field mean; faces: their slice
  double* w = (double*)petch_scratch(ctx, SCR_RAD_A, (size_t)(14 * (size_t)nz + 1) * sizeof(double));
  int* snap = (int*)petch_scratch(ctx, SCR_RAD_B, (size_t)F * sizeof(int));
  if (!w || !snap) return PETCH_E_HIP;
  double *area = w, *perim = w + nz, *loss = w + 2 * (size_t)nz, *prof = w + 4 * (size_t)nz,
         *ws = w + 6 * (size_t)nz, *fmean = w + 14 * (size_t)nz;
  const int trench = p->trench ? 1 : 0;
  hipLaunchKernelGGL(k_kn_slices, dim3((unsigned)nz), dim3(kBS), 0, ctx->stream, phi, nx, ny, nz,
                     p->dx, p->sub_top, trench, area, perim, loss);
  PETCH_LAUNCH_CHECK(ctx);
  if (p->sink_field) {
    hipLaunchKernelGGL(k_kn_field_mean, dim3(1), dim3(kBS), 0, ctx->stream, centroids, bare, F,
                       p->sub_top, p->dx, fmean);
    PETCH_LAUNCH_CHECK(ctx);
  }
  KnFace kf;
  kf.dx = p->dx; kf.sub_top = p->sub_top; kf.cx = p->cx; kf.cy = p->cy;
  kf.band = p->half_width + 2.0 * p->dx;
  kf.beta[0] = p->beta_a; kf.beta[1] = p->beta_b;
  kf.nz = nz; kf.trench = trench; kf.footprint = p->footprint ? 1 : 0;
  kf.sink_field = p->sink_field ? 1 : 0;
  hipLaunchKernelGGL(k_kn_snap, dim3(petch_blocks(F, kBS)), dim3(kBS), 0, ctx->stream, centroids,
                     gas_normals, areas, bare, F, kf, area, perim, fmean, snap, loss);
  PETCH_LAUNCH_CHECK(ctx);
  KnSolve ks;
  ks.dx = p->dx; ks.sub_top = p->sub_top; ks.Ly = p->Ly; ks.wls = p->wall_loss_scale;
  ks.band_W = p->band_W; ks.pass_frac = p->pass_frac; ks.ar_pass = p->ar_pass;
  ks.W = p->footprint ? 2.0 * p->half_width : p->dx;
  ks.nz = nz; ks.trench = trench; ks.front_loss = p->front_loss ? 1 : 0;
  hipLaunchKernelGGL(k_kn_solve, dim3(1), dim3(64), 0, ctx->stream, area, perim, loss, ks, prof, ws);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_kn_face, dim3(petch_blocks(F, kBS)), dim3(kBS), 0, ctx->stream, centroids,
                     snap, F, nz, p->sub_top, prof, m_a, m_b);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_dda_neutral_flux(petch_ctx* ctx, const float* phi, int nx, int ny, int nz,
                                      const float* centroids, const float* gas_normals,
                                      const float* areas, const float* bare, int F,
                                      const petch_dda_params* p, float* m_a, float* m_b) {
  PETCH_ARG(ctx, ctx && phi && p && m_a, "dda_neutral_flux: null pointer");
  PETCH_ARG(ctx, nx > 1 && ny > 1 && nz > 1 && F >= 0, "dda_neutral_flux: bad shape");
  PETCH_ARG(ctx, p->dx > 0.0 && p->n_dir > 0, "dda_neutral_flux: dx <= 0 or n_dir <= 0");
  if (F == 0) return PETCH_OK;
  PETCH_ARG(ctx, centroids && gas_normals && areas && bare, "dda_neutral_flux: null face array");
  const long long N = (long long)nx * ny * nz;
  const int nd = 2 * p->n_dir;
  // the direction set (fib_sphere(2 n_dir), dda.py:124-135): float64, handed over as float32
  std::vector<float> dirs((size_t)nd * 3);
  const double gold = M_PI * (1.0 + sqrt(5.0));
  for (int i = 0; i < nd; ++i) {
    const double t = (double)i + 0.5;
    const double z = 1.0 - 2.0 * t / (double)nd;
    const double r = sqrt(fmax(1.0 - z * z, 0.0));
    const double th = gold * t;
    dirs[3 * i] = (float)(r * cos(th));
    dirs[3 * i + 1] = (float)(r * sin(th));
    dirs[3 * i + 2] = (float)z;
  }
  // cells: area (double), re-emission sum (double), field (float), seeded (u8);
  // faces: cell, direct and gathered values (float), three iterates per species (double)
  char* cw = (char*)petch_scratch(ctx, SCR_RAD_A, (size_t)N * (8 + 8 + 4 + 1) + 64);
  char* fw = (char*)petch_scratch(ctx, SCR_RAD_B, (size_t)F * (4 + 4 + 4 + 6 * 8) + 64);
  if (!cw || !fw) return PETCH_E_HIP;
  double* carea = (double*)cw;
  double* R = carea + N;
  float* field = (float*)(R + N);
  uint8_t* seeded = (uint8_t*)(field + N);
  int* cell = (int*)fw;
  float* direct = (float*)(cell + F);
  float* gath = direct + F;
  double* it = (double*)(((uintptr_t)(gath + F) + 7) & ~(uintptr_t)7);
  const float* dd = (const float*)petch_upload_table(ctx, dirs.data(), dirs.size() * sizeof(float));
  if (!dd) return PETCH_E_HIP;
  const double dx = p->dx;
  const double z_top = (double)(nz - 1) * dx - 0.5 * dx;
  DdaGrid g;
  g.phi = phi; g.nx = nx; g.ny = ny; g.nz = nz;
  g.dx = (float)dx;
  g.z_top = (float)z_top;
  g.Ly = (float)ny * (float)dx;
  g.max_steps = ((int)(z_top / dx) + 4) * (p->periodic_y ? 3 : 1);
  g.periodic_y = p->periodic_y ? 1 : 0;
  const unsigned gN = grid_of(N), gF = petch_blocks(F, kBS), gW = petch_blocks(F, kBS / 64);
  hipLaunchKernelGGL(k_dda_zero, dim3(gN), dim3(kBS), 0, ctx->stream, carea, (double*)nullptr,
                     seeded, N);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_dda_cells, dim3(gF), dim3(kBS), 0, ctx->stream, centroids, gas_normals,
                     areas, F, nx, ny, nz, dx, cell, carea, seeded);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_dda_gather, dim3(gW), dim3(kBS), 0, ctx->stream, g, centroids, gas_normals,
                     F, dd, nd, (const float*)nullptr, 1.0f, direct);
  PETCH_LAUNCH_CHECK(ctx);
  const int n_re = p->n_reemit > 1 ? p->n_reemit : 1;
  for (int s = 0; s < 2; ++s) {
    float* out = s ? m_b : m_a;
    if (!out) continue;
    const double beta = s ? p->beta_b : p->beta_a;
    double* v[3] = {it + (size_t)3 * s * F, it + (size_t)(3 * s + 1) * F, it + (size_t)(3 * s + 2) * F};
    // v[0]: current iterate; v[1], v[2]: the two before it
    hipLaunchKernelGGL(k_dda_update, dim3(gF), dim3(kBS), 0, ctx->stream, direct,
                       (const float*)nullptr, F, v[0]);
    PETCH_LAUNCH_CHECK(ctx);
    int have = 1;
    for (int r = 0; r < n_re; ++r) {
      hipLaunchKernelGGL(k_dda_zero, dim3(gN), dim3(kBS), 0, ctx->stream, R, (double*)nullptr,
                         (uint8_t*)nullptr, N);
      PETCH_LAUNCH_CHECK(ctx);
      hipLaunchKernelGGL(k_dda_scatter, dim3(gF), dim3(kBS), 0, ctx->stream, cell, areas, bare,
                         beta, v[0], F, R);
      PETCH_LAUNCH_CHECK(ctx);
      hipLaunchKernelGGL(k_dda_mean, dim3(gN), dim3(kBS), 0, ctx->stream, R, carea, N);
      PETCH_LAUNCH_CHECK(ctx);
      hipLaunchKernelGGL(k_dda_fill, dim3(gN), dim3(kBS), 0, ctx->stream, phi, R, seeded, nx, ny,
                         nz, field);
      PETCH_LAUNCH_CHECK(ctx);
      hipLaunchKernelGGL(k_dda_gather, dim3(gW), dim3(kBS), 0, ctx->stream, g, centroids,
                         gas_normals, F, dd, nd, field, 0.0f, gath);
      PETCH_LAUNCH_CHECK(ctx);
      double* t = v[2];
      v[2] = v[1];
      v[1] = v[0];
      v[0] = t;
      hipLaunchKernelGGL(k_dda_update, dim3(gF), dim3(kBS), 0, ctx->stream, direct, gath, F, v[0]);
      PETCH_LAUNCH_CHECK(ctx);
      ++have;
    }
    hipLaunchKernelGGL(k_dda_final, dim3(gF), dim3(kBS), 0, ctx->stream, v[0], v[1],
                       have >= 3 ? v[2] : (double*)nullptr, F, out);
    PETCH_LAUNCH_CHECK(ctx);
  }
  return PETCH_OK;
}
