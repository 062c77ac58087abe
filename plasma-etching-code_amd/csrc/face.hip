// Per-face and per-node kernels between the traces and the level set:
//   petch_normalize       m = clamp((fl/max(A,1e-9))*scale, lo, hi)      threed.py:165-170,814-821
//   petch_cos_ratio       cos_i = ang/fl where fl > 0 else 0            threed.py:179-185,821
//   petch_smooth          edge-neighbour normal-weighted smoothing       threed.py:129-148,195-214
//   petch_normalize_adjoint, petch_smooth_adjoint: their transposes (no reference counterpart)
//   petch_belen_coverage  Belen coverages -> bare = clip(1-thF-thO,0,1)  threed.py:715-737,903-904
//   petch_belen_coverage_adjoint  its closed-form derivative (no reference counterpart)
//   petch_surface_rate    V per face (belen / langmuir), mask faces 0,  chemistry.py:62-92,
//                         non-finite -> 0 (counted), running max of V   belen.py:26-63,
//                                                                       threed.py:1608-1616,1721-1728,1750
//   petch_extend          Fs[node] = V[nearest face] in |phi| < band     threed.py:1324-1346
//   petch_surface_rate_adjoint, petch_extend_adjoint: their adjoints (no reference counterpart)
//   petch_nearest_face_value  out[i] = src[nearest face of pts[i]]       threed.py:1619-1639
//   petch_extend_batch, petch_nearest_face_value_batch: the same for many grids / point sets in
//                         one pass (the batched etch loop)
// The chemistry is evaluated in float64 (the reference evaluates it on the host in float64) and
// rounded to float32 once.
#include "common.h"
#include "bvh.h"
#include "fastdiv.h"
#include "scan.h"

#include <vector>

petch::BvhView petch_bvh_view(const petch_bvh* b);

namespace {

constexpr int kBS = 256;

__device__ __forceinline__ float norm_value(float fl, float area, float scale, float lo,
                                            float hi) {
  const float A = fmaxf(area, 1.0e-9f);
  return fminf(fmaxf((fl / A) * scale, lo), hi);
}

__global__ void k_normalize(const float* __restrict__ fl, const float* __restrict__ area, int F,
                            float scale, float lo, float hi, float* __restrict__ out) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  out[f] = norm_value(fl[f], area[f], scale, lo, hi);
}

__global__ void k_cos_ratio(const float* __restrict__ fl, const float* __restrict__ ang, int F,
                            float* __restrict__ out) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  float x = fl[f];
  out[f] = x > 0.f ? ang[f] / x : 0.f;
}

// One smoothing pass: out_i = (x_i + sum_j w_ij x_j) / (1 + sum_j w_ij), w_ij = max(0, n_i.n_j),
// j over the (up to 3) edge neighbours; blended out = (1-alpha)*raw + alpha*out when alpha < 1.
// The smoothing pass at face f over the values val(g) of f and its neighbours (a functor, so a
// fused caller can compute them on the fly).
template <class V>
__device__ __forceinline__ float smooth_value(const V& val, const int* __restrict__ nbr,
                                              const float* __restrict__ nrm, int F, float alpha,
                                              float raw, int f) {
  float nx = nrm[3 * f], ny = nrm[3 * f + 1], nz = nrm[3 * f + 2];
  float num = val(f), den = 1.f;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    int g = nbr[3 * f + e];
    if (g < 0 || g >= F) continue;
    float w = fmaxf(0.f, nx * nrm[3 * g] + ny * nrm[3 * g + 1] + nz * nrm[3 * g + 2]);
    num += val(g) * w;
    den += w;
  }
  float s = num / fmaxf(den, 1.0e-12f);
  return (alpha < 1.f) ? (1.f - alpha) * raw + alpha * s : s;
}

__global__ void k_smooth(const float* __restrict__ x, const int* __restrict__ nbr,
                         const float* __restrict__ nrm, int F, float alpha,
                         const float* __restrict__ raw, float* __restrict__ out) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  auto val = [x](int g) { return x[g]; };
  out[f] = smooth_value(val, nbr, nrm, F, alpha, raw[f], f);
}

// Adjoint of k_normalize: gfl = gout * scale / A where the clamp is inactive (lo <= v <= hi).
__global__ void k_normalize_adjoint(const float* __restrict__ fl, const float* __restrict__ area,
                                    int F, float scale, float lo, float hi,
                                    const float* __restrict__ gout, float* __restrict__ gfl) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  float A = fmaxf(area[f], 1.0e-9f);
  float v = (fl[f] / A) * scale;
  gfl[f] = (v >= lo && v <= hi) ? gout[f] * (scale / A) : 0.f;
}

// The smoothing pass is s = D^-1 (I + W) x with W_fg = max(0, n_f.n_g) over edge neighbours and
// D_f = 1 + sum_g W_fg. Edge adjacency is symmetric and so is W, so the transposed pass is
// (I + W) D^-1 y: z = y / D (k_smooth_den), then out_g = a (z_g + sum_f W_gf z_f) + c y_g
// (k_smooth_transposed; c carries a per-pass blend with its own input).
__global__ void k_smooth_den(const float* __restrict__ y, const int* __restrict__ nbr,
                             const float* __restrict__ nrm, int F, float* __restrict__ z) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  float nx = nrm[3 * f], ny = nrm[3 * f + 1], nz = nrm[3 * f + 2];
  float den = 1.f;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    int g = nbr[3 * f + e];
    if (g < 0 || g >= F) continue;
    den += fmaxf(0.f, nx * nrm[3 * g] + ny * nrm[3 * g + 1] + nz * nrm[3 * g + 2]);
  }
  z[f] = y[f] / fmaxf(den, 1.0e-12f);
}

__global__ void k_smooth_transposed(const float* __restrict__ z, const float* y,
                                    const int* __restrict__ nbr, const float* __restrict__ nrm,
                                    int F, float a, float c, float* out) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  float nx = nrm[3 * f], ny = nrm[3 * f + 1], nz = nrm[3 * f + 2];
  float t = z[f];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    int g = nbr[3 * f + e];
    if (g < 0 || g >= F) continue;
    t += z[g] * fmaxf(0.f, nx * nrm[3 * g] + ny * nrm[3 * g + 1] + nz * nrm[3 * g + 2]);
  }
  float yf = y[f];  // y may alias out: each lane reads its own entry before writing it
  out[f] = a * t + c * yf;
}

__global__ void k_axpy(float* __restrict__ y, const float* __restrict__ x, float a, int F) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f < F) y[f] += a * x[f];
}

struct Chem {
  double ionFlux, Fflux, Oflux, k_sigma, beta_sigma, rho, rate_scale, B_sp, fnE, fnO;
  double Yie, Ysp, Yp, s_F, s_O;
  int viennaps, belen;
};

__device__ __forceinline__ void angular(double cos_i, const Chem& c, double& f_ie, double& f_sp) {
  double x = fmin(fmax(cos_i, 0.0), 1.0);
  if (c.viennaps) {
    f_sp = fmax((1.0 + c.B_sp * (1.0 - x * x)) * x, 0.0);
    double theta = acos(x);
    f_ie = (x >= 0.5) ? 1.0 : fmax(3.0 - 6.0 * theta / 3.141592653589793, 0.0);
  } else {
    f_ie = x;
    f_sp = x;
  }
}

// Belen coverages of face f for the normalised fluxes mF, mO: bare = clip(1 - thF - thO, 0, 1).
__device__ __forceinline__ float belen_value(const float* __restrict__ m_i, float mF, float mO,
                                             const float* __restrict__ cos_i,
                                             const float* __restrict__ iy_ie,
                                             const float* __restrict__ iy_ip, const Chem& c, int f,
                                             double& thF, double& thO) {
  const double eps = 1e-9;
  double GY_ie, GY_p;
  if (iy_ie) {
    GY_ie = c.ionFlux * (double)iy_ie[f];
    GY_p = c.ionFlux * (double)iy_ip[f];
  } else {
    double f_ie, f_sp;
    angular((double)cos_i[f], c, f_ie, f_sp);
    double Fi = c.ionFlux * (double)m_i[f];
    GY_ie = c.Yie * f_ie * Fi;
    GY_p = c.Yp * f_ie * Fi;
  }
  double GbE = c.Fflux * (double)mF + eps;
  double GbP = c.Oflux * (double)mO + eps;
  double a = (c.k_sigma + 2.0 * GY_ie) / GbE;
  double b = (c.beta_sigma + GY_p) / GbP;
  thF = 1.0 / (1.0 + a * (1.0 + 1.0 / (b + eps)));
  thO = 1.0 / (1.0 + b * (1.0 + 1.0 / (a + eps)));
  double br = 1.0 - thF - thO;
  return (float)fmin(fmax(br, 0.0), 1.0);
}

__global__ void k_belen_coverage(const float* __restrict__ m_i, const float* __restrict__ m_F,
                                 const float* __restrict__ m_O, const float* __restrict__ cos_i,
                                 const float* __restrict__ iy_ie, const float* __restrict__ iy_ip,
                                 Chem c, int F, float* __restrict__ bare,
                                 float* __restrict__ thF_out, float* __restrict__ thO_out) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  double thF, thO;
  bare[f] = belen_value(m_i, m_F[f], m_O[f], cos_i, iy_ie, iy_ip, c, f, thF, thO);
  if (thF_out) thF_out[f] = (float)thF;
  if (thO_out) thO_out[f] = (float)thO;
}

// The end of one neutral fixed-point round of many features (petch.batch): per face, both
// species' accumulators normalised and smoothed once (normalised neighbour values computed on
// the fly), then the Belen update; the arithmetic of k_normalize, k_smooth and
// k_belen_coverage, shared through the same device functions. Feature t owns the workgroups
// [blk0[t], blk0[t+1]).
struct RoundItem {
  const float *flF, *flO, *area, *nrm;
  const int* nbr;
  const float *m_i, *cos_i, *iy_ie, *iy_ip;
  float *m_F, *m_O, *bare;
  Chem c;
  int nf, smooth;
  float scale, hi, alpha;
  long long blk0;
};

__global__ void k_finish_round(const RoundItem* __restrict__ items, int n) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].blk0 <= (long long)blockIdx.x) lo = mid; else hi = mid - 1;
  }
  const RoundItem& it = items[lo];
  const int f = (int)(blockIdx.x - it.blk0) * kBS + threadIdx.x;
  if (f >= it.nf) return;
  const float* area = it.area;
  const float scale = it.scale, hv = it.hi;
  const float* flF = it.flF;
  const float* flO = it.flO;
  auto vF = [=](int g) { return norm_value(flF[g], area[g], scale, 0.f, hv); };
  auto vO = [=](int g) { return norm_value(flO[g], area[g], scale, 0.f, hv); };
  float mF = vF(f), mO = vO(f);
  if (it.smooth) {
    mF = smooth_value(vF, it.nbr, it.nrm, it.nf, it.alpha, mF, f);
    mO = smooth_value(vO, it.nbr, it.nrm, it.nf, it.alpha, mO, f);
  }
  it.m_F[f] = mF;
  it.m_O[f] = mO;
  double thF, thO;
  it.bare[f] = belen_value(it.m_i, mF, mO, it.cos_i, it.iy_ie, it.iy_ip, it.c, f, thF, thO);
}

// The end of the faithful-ion trace of many features (petch.batch): the three yield channels and
// the arrival count normalised without a clip and smoothed once (k_normalize, k_smooth), and
// cos_i = ang / arrivals (k_cos_ratio), through the same device functions. acc is the trace's
// (5, nf) accumulator (sp, ie, ip, arrivals, angle sum); out receives (sp, ie, ip, m_i, cos_i)
// as five rows of nf.
struct IonItem {
  const float *acc, *area, *nrm;
  const int* nbr;
  float* out;
  int nf, smooth;
  float scale, alpha;
  long long blk0;
};

__global__ void k_finish_ion(const IonItem* __restrict__ items, int n) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].blk0 <= (long long)blockIdx.x) lo = mid; else hi = mid - 1;
  }
  const IonItem& it = items[lo];
  const int f = (int)(blockIdx.x - it.blk0) * kBS + threadIdx.x;
  const int F = it.nf;
  if (f >= F) return;
  const float* area = it.area;
  const float scale = it.scale;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float* a = it.acc + (long long)k * F;
    auto v = [=](int g) { return norm_value(a[g], area[g], scale, -INFINITY, INFINITY); };
    float m = v(f);
    if (it.smooth) m = smooth_value(v, it.nbr, it.nrm, F, it.alpha, m, f);
    it.out[(long long)k * F + f] = m;
  }
  const float x = it.acc[3LL * F + f];
  it.out[4LL * F + f] = x > 0.f ? it.acc[4LL * F + f] / x : 0.f;
}

// Closed-form adjoint of k_belen_coverage for gbare = d loss / d bare (the clip passes the
// gradient where 0 <= 1 - thF - thO <= 1). With a = (k_sigma + 2 GY_ie) / GbE and
// b = (beta_sigma + GY_p) / GbP:
//   dthF/da = -(1 + 1/(b+eps)) thF^2   dthF/db = a thF^2 / (b+eps)^2
//   dthO/db = -(1 + 1/(a+eps)) thO^2   dthO/da = b thO^2 / (a+eps)^2
// and da/dm_F = -a Fflux / GbE, db/dm_O = -b Oflux / GbP, da/dGY_ie = 2 / GbE, db/dGY_p = 1 / GbP.
// Outputs d loss / d m_F and d m_O; with ion yields also d / d iy_ie and d iy_ip (g_ie, g_ip), else
// d / d m_i through GY = Y f_ie ionFlux m_i (gm_i); cos_i is held constant. Optional outputs may
// be NULL.
__global__ void k_belen_coverage_adjoint(
    const float* __restrict__ m_i, const float* __restrict__ m_F, const float* __restrict__ m_O,
    const float* __restrict__ cos_i, const float* __restrict__ iy_ie,
    const float* __restrict__ iy_ip, Chem c, int F, const float* __restrict__ gbare,
    float* __restrict__ gm_F, float* __restrict__ gm_O, float* __restrict__ gm_i,
    float* __restrict__ g_ie, float* __restrict__ g_ip) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  const double eps = 1e-9;
  double GY_ie, GY_p, f_ie = 0.0;
  if (iy_ie) {
    GY_ie = c.ionFlux * (double)iy_ie[f];
    GY_p = c.ionFlux * (double)iy_ip[f];
  } else {
    double f_sp;
    angular((double)cos_i[f], c, f_ie, f_sp);
    double Fi = c.ionFlux * (double)m_i[f];
    GY_ie = c.Yie * f_ie * Fi;
    GY_p = c.Yp * f_ie * Fi;
  }
  double GbE = c.Fflux * (double)m_F[f] + eps;
  double GbP = c.Oflux * (double)m_O[f] + eps;
  double a = (c.k_sigma + 2.0 * GY_ie) / GbE;
  double b = (c.beta_sigma + GY_p) / GbP;
  double ib = 1.0 / (b + eps), ia = 1.0 / (a + eps);
  double thF = 1.0 / (1.0 + a * (1.0 + ib));
  double thO = 1.0 / (1.0 + b * (1.0 + ia));
  double br = 1.0 - thF - thO;
  double g = (br >= 0.0 && br <= 1.0) ? (double)gbare[f] : 0.0;
  // bare = 1 - thF - thO
  double ga = -g * (-(1.0 + ib) * thF * thF + b * thO * thO * ia * ia);
  double gb = -g * (a * thF * thF * ib * ib - (1.0 + ia) * thO * thO);
  gm_F[f] = (float)(ga * (-a * c.Fflux / GbE));
  gm_O[f] = (float)(gb * (-b * c.Oflux / GbP));
  double gGie = ga * 2.0 / GbE, gGp = gb / GbP;
  if (iy_ie) {
    if (g_ie) g_ie[f] = (float)(gGie * c.ionFlux);
    if (g_ip) g_ip[f] = (float)(gGp * c.ionFlux);
  } else if (gm_i) {
    gm_i[f] = (float)((gGie * c.Yie + gGp * c.Yp) * f_ie * c.ionFlux);
  }
}

struct MaskGeom {
  double sub_top, top, cx, cy, hw, hw2;
  int hole;
};

// faces_in_mask on the centroid (threed.py:1608-1616), or the given mask
__device__ __forceinline__ bool face_masked(const uint8_t* __restrict__ is_mask,
                                            const float* __restrict__ centroids,
                                            const MaskGeom& mg, int f) {
  if (is_mask) return is_mask[f] != 0;
  double x = centroids[3 * f], y = centroids[3 * f + 1], z = centroids[3 * f + 2];
  bool in_band = (z >= mg.sub_top) && (z < mg.top);
  bool opening;
  if (mg.hole) {
    double ax = x - mg.cx, ay = y - mg.cy;
    opening = ax * ax + ay * ay < mg.hw2;
  } else {
    opening = fabs(x - mg.cx) < mg.hw;
  }
  return in_band && !opening;
}

__device__ __forceinline__ void rate_face(const float* __restrict__ m_i,
                                          const float* __restrict__ m_F,
                                          const float* __restrict__ m_O,
                                          const float* __restrict__ cos_i,
                                          const float* __restrict__ iy_sp,
                                          const float* __restrict__ iy_ie,
                                          const float* __restrict__ iy_ip,
                                          const float* __restrict__ centroids,
                                          const uint8_t* __restrict__ is_mask, const MaskGeom& mg,
                                          const Chem& c, int f, float* __restrict__ V,
                                          unsigned& vbits, unsigned& bad) {
#pragma clang fp contract(off)
  const double eps = 1e-9;
  double f_ie, f_sp;
  angular((double)cos_i[f], c, f_ie, f_sp);
  double Fi = c.ionFlux * (double)m_i[f];
  double Yie_a = c.Yie * f_ie, Ysp_a = c.Ysp * f_sp, Yp_a = c.Yp * f_ie;
  double v;
  if (c.belen) {
    double GY_ie, GY_p, sputter, ion_enh;
    if (iy_ie) {
      GY_ie = c.ionFlux * (double)iy_ie[f];
      GY_p = c.ionFlux * (double)iy_ip[f];
      sputter = c.ionFlux * (double)iy_sp[f];
      ion_enh = GY_ie;
    } else {
      GY_ie = Yie_a * Fi;
      GY_p = Yp_a * Fi;
      sputter = Ysp_a * Fi;
      ion_enh = Yie_a * Fi;
    }
    double GbE = c.Fflux * (double)m_F[f] / c.fnE + eps;
    double GbP = c.Oflux * (double)m_O[f] / c.fnO + eps;
    double a = (c.k_sigma + 2.0 * GY_ie) / GbE;
    double b = (c.beta_sigma + GY_p) / GbP;
    double thF = 1.0 / (1.0 + a * (1.0 + 1.0 / (b + eps)));
    double rate = c.k_sigma * thF / 4.0 + sputter + thF * ion_enh;
    v = (1.0 / c.rho) * rate * c.rate_scale;
  } else {
    double Fev = c.Fflux * (double)m_F[f], Fp = c.Oflux * (double)m_O[f];
    double rF = c.s_F * Fev / (Yie_a * Fi + eps);
    double rO = c.s_O * Fp / (Yp_a * Fi + eps);
    double br = 1.0 / (1.0 + rF + rO);
    double thF = rF * br;
    v = (1.0 / c.rho) * (Yie_a * Fi * thF + Ysp_a * Fi * br) * c.rate_scale;
  }
  if (face_masked(is_mask, centroids, mg, f)) v = 0.0;
  float vf = (float)v;
  if (!(fabsf(vf) <= 3.402823466e38f)) {  // NaN or inf
    vf = 0.f;
    bad = 1u;
  }
  V[f] = vf;
  // V >= 0 here, so the float bit pattern orders like an unsigned integer
  vbits = __float_as_uint(vf > 0.f ? vf : 0.f);
}

// red[0] = max(red[0], the wavefront's largest V bits), red[1] += its non-finite count: one
// atomic each per wavefront (all its lanes belong to one surface; lanes without a face pass 0).
// Max and count do not depend on the order, so the pair is the same as with per-face atomics.
__device__ __forceinline__ void rate_reduce(unsigned* __restrict__ red, unsigned vbits,
                                            unsigned bad) {
  unsigned m = vbits;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, d, 64));
  const unsigned long long b = __ballot(bad != 0u);
  if ((threadIdx.x & 63) == 0) {
    if (m) atomicMax(red, m);
    if (b) atomicAdd(red + 1, (unsigned)__popcll(b));
  }
}

__global__ void k_surface_rate(const float* __restrict__ m_i, const float* __restrict__ m_F,
                               const float* __restrict__ m_O, const float* __restrict__ cos_i,
                               const float* __restrict__ iy_sp, const float* __restrict__ iy_ie,
                               const float* __restrict__ iy_ip,
                               const float* __restrict__ centroids,
                               const uint8_t* __restrict__ is_mask, MaskGeom mg, Chem c, int F,
                               float* __restrict__ V, unsigned* __restrict__ red) {
  int f = blockIdx.x * kBS + threadIdx.x;
  unsigned vbits = 0u, bad = 0u;
  if (f < F)
    rate_face(m_i, m_F, m_O, cos_i, iy_sp, iy_ie, iy_ip, centroids, is_mask, mg, c, f, V, vbits,
              bad);
  rate_reduce(red, vbits, bad);
}

// The surface rates of many surfaces in one launch: item t owns the workgroups
// [blk0, blk0 + ceil(F / kBS)) (none for an empty surface), and its reduction pair is cleared,
// and its BVH guard flag copied, by k_rate_clear before.
struct RateItem {
  const float *m_i, *m_F, *m_O, *cos_i, *iy_sp, *iy_ie, *iy_ip, *centroids;
  const uint8_t* is_mask;
  MaskGeom mg;
  Chem c;
  int F;
  float* V;
  unsigned* red;
  const int* err_src;
  int* err_dst;
  long long blk0;
};

__global__ void k_rate_clear(const RateItem* __restrict__ it, int n) {
  const int t = blockIdx.x * kBS + threadIdx.x;
  if (t >= n) return;
  it[t].red[0] = 0u;
  it[t].red[1] = 0u;
  if (it[t].err_dst) *it[t].err_dst = it[t].err_src ? *it[t].err_src : 0;
}

__global__ void k_rate_many(const RateItem* __restrict__ items, int n) {
  int lo = 0, hi = n - 1;
  const long long b = blockIdx.x;
  while (lo < hi) {                // the last item starting at or before block b owns it
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].blk0 <= b) lo = mid; else hi = mid - 1;
  }
  const RateItem& it = items[lo];
  const int f = (int)(b - it.blk0) * kBS + threadIdx.x;
  unsigned vbits = 0u, bad = 0u;
  if (f < it.F)
    rate_face(it.m_i, it.m_F, it.m_O, it.cos_i, it.iy_sp, it.iy_ie, it.iy_ip, it.centroids,
              it.is_mask, it.mg, it.c, f, it.V, vbits, bad);
  rate_reduce(it.red, vbits, bad);
}

// Closed-form adjoint of k_surface_rate for gV = d loss / dV (zero on mask faces and where V
// was not finite): gm_F, gm_O, and d / d m_i (gm_i) or, with ion yields, d / d iy_sp, iy_ie,
// iy_ip (g_sp, g_ie, g_ip); cos_i is held constant. gY (3 x F) receives each face's share of
// d loss / d (Yie, Ysp, Yp), the yield coefficients of the analytic-yield path (zero with ion
// yields), whose face sums give the gradient in the recipe (ion energy, yield prefactors).
// Optional outputs may be NULL.
//   belen:    V = K (k_sigma thF / 4 + sputter + thF ion_enh), thF = 1 / (1 + a (1 + 1/(b+eps)))
//   langmuir: V = K (Yie_a Fi thF + Ysp_a Fi br), br = 1 / (1 + rF + rO), thF = rF br
// with K = rate_scale / rho.
__global__ void k_surface_rate_adjoint(
    const float* __restrict__ m_i, const float* __restrict__ m_F, const float* __restrict__ m_O,
    const float* __restrict__ cos_i, const float* __restrict__ iy_sp,
    const float* __restrict__ iy_ie, const float* __restrict__ iy_ip,
    const float* __restrict__ centroids, const uint8_t* __restrict__ is_mask, MaskGeom mg, Chem c,
    int F, const float* __restrict__ gV, float* __restrict__ gm_F, float* __restrict__ gm_O,
    float* __restrict__ gm_i, float* __restrict__ g_sp, float* __restrict__ g_ie,
    float* __restrict__ g_ip, float* __restrict__ gY) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= F) return;
  const double eps = 1e-9;
  double yie = 0.0, ysp = 0.0, yp = 0.0;  // dV / dYie, dYsp, dYp (analytic-yield path)
  double f_ie, f_sp;
  angular((double)cos_i[f], c, f_ie, f_sp);
  const double Fi = c.ionFlux * (double)m_i[f];
  const double Yie_a = c.Yie * f_ie, Ysp_a = c.Ysp * f_sp, Yp_a = c.Yp * f_ie;
  const double K = c.rate_scale / c.rho;
  double v, dF = 0.0, dO = 0.0, dI = 0.0, dsp = 0.0, die = 0.0, dip = 0.0;
  if (c.belen) {
    double GY_ie, GY_p, sputter, ion_enh;
    if (iy_ie) {
      GY_ie = c.ionFlux * (double)iy_ie[f];
      GY_p = c.ionFlux * (double)iy_ip[f];
      sputter = c.ionFlux * (double)iy_sp[f];
      ion_enh = GY_ie;
    } else {
      GY_ie = Yie_a * Fi;
      GY_p = Yp_a * Fi;
      sputter = Ysp_a * Fi;
      ion_enh = Yie_a * Fi;
    }
    const double GbE = c.Fflux * (double)m_F[f] / c.fnE + eps;
    const double GbP = c.Oflux * (double)m_O[f] / c.fnO + eps;
    const double a = (c.k_sigma + 2.0 * GY_ie) / GbE;
    const double b = (c.beta_sigma + GY_p) / GbP;
    const double ib = 1.0 / (b + eps);
    const double thF = 1.0 / (1.0 + a * (1.0 + ib));
    v = K * (c.k_sigma * thF / 4.0 + sputter + thF * ion_enh);
    const double dth = K * (c.k_sigma / 4.0 + ion_enh);   // dV / dthF
    const double ta = -(1.0 + ib) * thF * thF;             // dthF / da
    const double tb = a * thF * thF * ib * ib;             // dthF / db
    dF = dth * ta * (-a * (c.Fflux / c.fnE) / GbE);
    dO = dth * tb * (-b * (c.Oflux / c.fnO) / GbP);
    const double dGie = dth * ta * 2.0 / GbE, dGp = dth * tb / GbP;  // dV / dGY_ie, dGY_p
    if (iy_ie) {
      dsp = K * c.ionFlux;
      die = (dGie + K * thF) * c.ionFlux;
      dip = dGp * c.ionFlux;
    } else {
      dI = (dGie * Yie_a + dGp * Yp_a + K * Ysp_a + K * thF * Yie_a) * c.ionFlux;
      yie = (dGie + K * thF) * f_ie * Fi;
      ysp = K * f_sp * Fi;
      yp = dGp * f_ie * Fi;
    }
  } else {
    const double Fev = c.Fflux * (double)m_F[f], Fp = c.Oflux * (double)m_O[f];
    const double dE = Yie_a * Fi + eps, dP = Yp_a * Fi + eps;
    const double rF = c.s_F * Fev / dE, rO = c.s_O * Fp / dP;
    const double br = 1.0 / (1.0 + rF + rO);
    const double thF = rF * br;
    v = K * (Yie_a * Fi * thF + Ysp_a * Fi * br);
    // dV/drF and dV/drO at fixed Fi (dbr/dr = -br^2, dthF/drF = br - rF br^2, dthF/drO = -rF br^2)
    const double vF = K * (Yie_a * Fi * (br - rF * br * br) - Ysp_a * Fi * br * br);
    const double vO = K * (-Yie_a * Fi * rF * br * br - Ysp_a * Fi * br * br);
    dF = vF * c.s_F * c.Fflux / dE;
    dO = vO * c.s_O * c.Oflux / dP;
    // m_i enters through Fi: explicitly, and through rF, rO (drF/dFi = -rF Yie_a / dE)
    const double dFi = K * (Yie_a * thF + Ysp_a * br) + vF * (-rF * Yie_a / dE) +
                       vO * (-rO * Yp_a / dP);
    dI = dFi * c.ionFlux;
    yie = K * f_ie * Fi * thF + vF * (-rF * f_ie * Fi / dE);
    ysp = K * f_sp * Fi * br;
    yp = vO * (-rO * f_ie * Fi / dP);
  }
  double g = (double)gV[f];
  if (face_masked(is_mask, centroids, mg, f) || !(fabs((double)(float)v) <= 3.402823466e38))
    g = 0.0;
  gm_F[f] = (float)(g * dF);
  gm_O[f] = (float)(g * dO);
  if (iy_ie && c.belen) {
    if (g_sp) g_sp[f] = (float)(g * dsp);
    if (g_ie) g_ie[f] = (float)(g * die);
    if (g_ip) g_ip[f] = (float)(g * dip);
  } else if (gm_i) {
    gm_i[f] = (float)(g * dI);
  }
  if (gY) {
    gY[f] = (float)(g * yie);
    gY[F + f] = (float)(g * ysp);
    gY[2 * F + f] = (float)(g * yp);
  }
}

// Velocity extension of one or many grids in one pass: flag every grid's band (and zero its
// Fs), scan the flags of all grids together, scatter the band nodes' global indices, then one
// nearest-face query per listed node, each on its own grid's surface. The queries are
// latency-bound chains of node fetches (a few tens of microseconds each), so a grid's
// extension alone leaves the GPU mostly idle; the queries of many grids in one launch overlap.
struct ExtendItem {
  petch::BvhView bv;
  const float* V;
  const float* phi;
  float* Fs;
  int ny, nz;
  petch::FastDiv dy, dz;
  double dx, band;
  long long end;  // exclusive end of this grid's nodes in the global node numbering
};

// grid coordinates of node m of an item (m < 2^31: the batch's node numbering is int32)
__device__ __forceinline__ petch::Vec3 item_node_pos(const ExtendItem& e, long long m) {
  const uint32_t u = (uint32_t)m;
  const uint32_t c = petch::fastdiv(u, e.dz);
  const uint32_t k = u - c * (uint32_t)e.nz;
  const uint32_t i = petch::fastdiv(c, e.dy);
  const uint32_t j = c - i * (uint32_t)e.ny;
  return petch::v3((float)((double)i * e.dx), (float)((double)j * e.dx), (float)((double)k * e.dx));
}

__device__ __forceinline__ int locate_item(const ExtendItem* __restrict__ it, int n, long long g,
                                           long long& local) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (it[mid].end > g) hi = mid; else lo = mid + 1;
  }
  local = g - (lo ? it[lo - 1].end : 0);
  return lo;
}

// The band |phi| < band of every item as the ascending list of its global node ids, by a
// two-pass compaction over tiles of kBandTile consecutive nodes: k_band_count counts each
// tile's band nodes (and, with zero_fs, sets Fs = 0 on every node, which the extension then
// fills on the band), one workgroup scans the tile counts into offsets (the band's size goes
// after the list), and k_band_emit writes each tile's band nodes in order at its offset.
// Ascending order keeps neighbouring queries on neighbouring nodes of one surface.
constexpr int kBandRounds = 8;
constexpr long long kBandTile = (long long)kBandRounds * kBS;

// A tile's nodes are found from the item of its first node, c0 (one search per workgroup, the
// same for all its lanes): a node moves on to later items only past their ends. A search per
// node (six dependent table loads each, eight nodes a lane) made the two passes latency-bound.
__device__ __forceinline__ bool band_node(const ExtendItem* __restrict__ it, int n, int c0,
                                          long long g, long long total, int zero_fs) {
  if (g >= total) return false;
  int c = c0;
  while (c < n - 1 && it[c].end <= g) ++c;
  const long long m = g - (c ? it[c - 1].end : 0);
  const ExtendItem& e = it[c];
  if (zero_fs) e.Fs[m] = 0.f;
  return fabs((double)e.phi[m]) < e.band;
}

__device__ __forceinline__ int tile_item(const ExtendItem* __restrict__ it, int n) {
  long long m;
  return locate_item(it, n, (long long)blockIdx.x * kBandTile, m);
}

__global__ void k_band_count(const ExtendItem* __restrict__ it, int n, long long total,
                             int* __restrict__ tile_cnt, int zero_fs) {
  const long long g0 = (long long)blockIdx.x * kBandTile + threadIdx.x;
  const int c0 = tile_item(it, n);
  int c = 0;
#pragma unroll
  for (int r = 0; r < kBandRounds; ++r)
    c += band_node(it, n, c0, g0 + (long long)r * kBS, total, zero_fs);
  const int s = petch::block_sum<kBS>(c);
  if (threadIdx.x == 0) tile_cnt[blockIdx.x] = s;
}

__global__ void k_band_emit(const ExtendItem* __restrict__ it, int n, long long total,
                            const int* __restrict__ tile_off, int* __restrict__ list) {
  const long long g0 = (long long)blockIdx.x * kBandTile + threadIdx.x;
  int base = tile_off[blockIdx.x];
  const int c0 = tile_item(it, n);
  for (int r = 0; r < kBandRounds; ++r) {
    const long long g = g0 + (long long)r * kBS;
    const bool in = band_node(it, n, c0, g, total, 0);
    int tot;
    const int ex = petch::block_excl_scan<kBS>(in ? 1 : 0, tot);
    if (in) list[base + ex] = (int)g;
    base += tot;
  }
}

// one nearest-face query per band node, over the list in grid-stride order
__global__ void k_extend(const ExtendItem* __restrict__ it, int n, const int* __restrict__ idx,
                         const int* __restrict__ count) {
  const int cnt = *count;
  for (int i = blockIdx.x * kBS + threadIdx.x; i < cnt; i += gridDim.x * kBS) {
    long long m;
    const ExtendItem& e = it[locate_item(it, n, idx[i], m)];
    const petch::PointHit h = petch::bvh_closest_point(e.bv, item_node_pos(e, m), 1.0e6f);
    e.Fs[m] = h.face >= 0 ? e.V[h.face] : 0.f;
  }
}

// Adjoint of k_extend (one grid): the extension copies V[nearest face] to each band node, so
// its transpose adds each band node's gFs (read through the item's Fs) to gV[nearest face].
__global__ void k_extend_adjoint(const ExtendItem* __restrict__ it, int n,
                                 const int* __restrict__ idx, const int* __restrict__ count,
                                 float* __restrict__ gV) {
  const int cnt = *count;
  for (int i = blockIdx.x * kBS + threadIdx.x; i < cnt; i += gridDim.x * kBS) {
    long long m;
    const ExtendItem& e = it[locate_item(it, n, idx[i], m)];
    const petch::PointHit h = petch::bvh_closest_point(e.bv, item_node_pos(e, m), 1.0e6f);
    if (h.face >= 0) atomicAdd(gV + h.face, e.Fs[m]);
  }
}

// Nearest-face values of many point sets, each on its own surface, in one launch.
struct NearestItem {
  petch::BvhView bv;
  const float* src;
  const float* pts;
  float* out;
  float miss;
  long long end;
};

__global__ void k_nearest_batch(const NearestItem* __restrict__ it, int n, long long total) {
  const long long g = (long long)blockIdx.x * kBS + threadIdx.x;
  if (g >= total) return;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (it[mid].end > g) hi = mid; else lo = mid + 1;
  }
  const NearestItem& e = it[lo];
  const long long i = g - (lo ? it[lo - 1].end : 0);
  const petch::PointHit h = petch::bvh_closest_point(
      e.bv, petch::v3(e.pts[3 * i], e.pts[3 * i + 1], e.pts[3 * i + 2]), 1.0e6f);
  e.out[i] = h.face >= 0 ? e.src[h.face] : e.miss;
}

Chem make_chem(const petch_chem_params* p) {
  Chem c;
  c.ionFlux = p->ionFlux; c.Fflux = p->Fflux; c.Oflux = p->Oflux; c.k_sigma = p->k_sigma;
  c.beta_sigma = p->beta_sigma; c.rho = p->rho; c.rate_scale = p->rate_scale; c.B_sp = p->B_sp;
  c.fnE = p->fnorm_E; c.fnO = p->fnorm_O; c.Yie = p->Yie; c.Ysp = p->Ysp; c.Yp = p->Yp;
  c.s_F = p->s_F; c.s_O = p->s_O; c.viennaps = p->angular_viennaps; c.belen = p->chemistry_belen;
  return c;
}

}  // namespace

extern "C" int petch_normalize(petch_ctx* ctx, const float* fl, const float* areas, int F,
                               float scale, float lo, float hi, float* out) {
  PETCH_ARG(ctx, ctx && fl && areas && out, "normalize: null pointer");
  PETCH_ARG(ctx, F >= 0, "normalize: F < 0");
  if (F == 0) return PETCH_OK;
  hipLaunchKernelGGL(k_normalize, dim3(petch_blocks(F, kBS)), dim3(kBS), 0, ctx->stream, fl, areas,
                     F, scale, lo, hi, out);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_normalize_adjoint(petch_ctx* ctx, const float* fl, const float* areas,
                                       int F, float scale, float lo, float hi, const float* gout,
                                       float* gfl) {
  PETCH_ARG(ctx, ctx && fl && areas && gout && gfl, "normalize_adjoint: null pointer");
  PETCH_ARG(ctx, F >= 0, "normalize_adjoint: F < 0");
  if (F == 0) return PETCH_OK;
  hipLaunchKernelGGL(k_normalize_adjoint, dim3(petch_blocks(F, kBS)), dim3(kBS), 0, ctx->stream,
                     fl, areas, F, scale, lo, hi, gout, gfl);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_cos_ratio(petch_ctx* ctx, const float* fl, const float* ang, int F,
                               float* out) {
  PETCH_ARG(ctx, ctx && fl && ang && out, "cos_ratio: null pointer");
  if (F <= 0) return PETCH_OK;
  hipLaunchKernelGGL(k_cos_ratio, dim3(petch_blocks(F, kBS)), dim3(kBS), 0, ctx->stream, fl, ang,
                     F, out);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_smooth(petch_ctx* ctx, const float* src, const int* nbr, const float* normals,
                            int F, int n_iter, float alpha, int blend_each, float* out) {
  PETCH_ARG(ctx, ctx && src && nbr && normals && out, "smooth: null pointer");
  PETCH_ARG(ctx, F >= 0, "smooth: F < 0");
  PETCH_ARG(ctx, src != out, "smooth: src and out must differ");
  if (F == 0) return PETCH_OK;
  if (n_iter <= 0 || alpha <= 0.f) {
    PETCH_HIP(ctx, hipMemcpyAsync(out, src, (size_t)F * 4, hipMemcpyDeviceToDevice, ctx->stream));
    return PETCH_OK;
  }
  float* tmp = (float*)petch_scratch(ctx, SCR_SMOOTH, (size_t)F * 4);
  if (!tmp) return PETCH_E_HIP;
  dim3 g(petch_blocks(F, kBS)), b(kBS);
  // ping-pong so the last pass lands in `out`
  const float* cur = src;
  for (int it = 0; it < n_iter; ++it) {
    float* dst = ((n_iter - 1 - it) % 2 == 0) ? out : tmp;
    bool last = it == n_iter - 1;
    // blend_each: every pass blends with its own input (device path, threed.py:195-214);
    // otherwise only the last pass blends, with the original input (threed.py:129-148)
    float a = blend_each ? alpha : (last ? alpha : 1.f);
    const float* raw = blend_each ? cur : src;
    hipLaunchKernelGGL(k_smooth, g, b, 0, ctx->stream, cur, nbr, normals, F, a, raw, dst);
    PETCH_LAUNCH_CHECK(ctx);
    cur = dst;
  }
  return PETCH_OK;
}

extern "C" int petch_smooth_adjoint(petch_ctx* ctx, const int* nbr, const float* normals, int F,
                                    int n_iter, float alpha, int blend_each, const float* gout,
                                    float* gsrc) {
  PETCH_ARG(ctx, ctx && nbr && normals && gout && gsrc, "smooth_adjoint: null pointer");
  PETCH_ARG(ctx, F >= 0, "smooth_adjoint: F < 0");
  PETCH_ARG(ctx, gout != gsrc, "smooth_adjoint: gout and gsrc must differ");
  if (F == 0) return PETCH_OK;
  if (n_iter <= 0 || alpha <= 0.f) {
    PETCH_HIP(ctx, hipMemcpyAsync(gsrc, gout, (size_t)F * 4, hipMemcpyDeviceToDevice, ctx->stream));
    return PETCH_OK;
  }
  float* z = (float*)petch_scratch(ctx, SCR_SMOOTH, (size_t)F * 4);
  if (!z) return PETCH_E_HIP;
  dim3 g(petch_blocks(F, kBS)), b(kBS);
  // the passes of petch_smooth in reverse; the first transposed pass reads gout, the others run
  // in place on gsrc
  const float* y = gout;
  for (int it = n_iter - 1; it >= 0; --it) {
    const bool last = it == n_iter - 1;
    const float a = blend_each ? alpha : (last ? alpha : 1.f);
    const bool blend = a < 1.f;
    const float aa = blend ? a : 1.f;
    const float c = (blend && blend_each) ? 1.f - a : 0.f;
    hipLaunchKernelGGL(k_smooth_den, g, b, 0, ctx->stream, y, nbr, normals, F, z);
    PETCH_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_smooth_transposed, g, b, 0, ctx->stream, z, y, nbr, normals, F, aa, c,
                       gsrc);
    PETCH_LAUNCH_CHECK(ctx);
    y = gsrc;
  }
  if (!blend_each && alpha < 1.f) {  // the final blend's (1 - alpha) src term
    hipLaunchKernelGGL(k_axpy, g, b, 0, ctx->stream, gsrc, gout, 1.f - alpha, F);
    PETCH_LAUNCH_CHECK(ctx);
  }
  return PETCH_OK;
}

extern "C" int petch_belen_coverage(petch_ctx* ctx, const float* m_i, const float* m_F,
                                    const float* m_O, const float* cos_i, const float* iy_ie,
                                    const float* iy_ip, const petch_chem_params* cp, int F,
                                    float* bare, float* thF, float* thO) {
  PETCH_ARG(ctx, ctx && m_F && m_O && cp && bare, "belen_coverage: null pointer");
  PETCH_ARG(ctx, (iy_ie && iy_ip) || (m_i && cos_i), "belen_coverage: need ion yields or m_i/cos_i");
  if (F <= 0) return PETCH_OK;
  hipLaunchKernelGGL(k_belen_coverage, dim3(petch_blocks(F, kBS)), dim3(kBS), 0, ctx->stream, m_i,
                     m_F, m_O, cos_i, iy_ie, iy_ip, make_chem(cp), F, bare, thF, thO);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_finish_round_batch(petch_ctx* ctx, int n, const petch_round_item* items) {
  PETCH_ARG(ctx, ctx && (n == 0 || items), "finish_round_batch: null pointer");
  PETCH_ARG(ctx, n >= 0, "finish_round_batch: n < 0");
  if (n == 0) return PETCH_OK;
  std::vector<RoundItem> t;
  t.reserve((size_t)n);
  long long blocks = 0;
  for (int k = 0; k < n; ++k) {
    const petch_round_item& q = items[k];
    PETCH_ARG(ctx, q.nf >= 0 && q.flF && q.flO && q.area && q.m_F && q.m_O && q.bare,
              "finish_round_batch: item %d is invalid", k);
    PETCH_ARG(ctx, !q.smooth || (q.nbr && q.normals), "finish_round_batch: item %d needs the "
              "adjacency and normals to smooth", k);
    PETCH_ARG(ctx, (q.iy_ie && q.iy_ip) || (q.m_i && q.cos_i),
              "finish_round_batch: item %d needs ion yields or m_i/cos_i", k);
    if (q.nf == 0) continue;
    RoundItem r;
    r.flF = q.flF; r.flO = q.flO; r.area = q.area; r.nrm = q.normals; r.nbr = q.nbr;
    r.m_i = q.m_i; r.cos_i = q.cos_i; r.iy_ie = q.iy_ie; r.iy_ip = q.iy_ip;
    r.m_F = q.m_F; r.m_O = q.m_O; r.bare = q.bare;
    r.c = make_chem(&q.chem);
    r.nf = q.nf;
    r.smooth = q.smooth && q.alpha > 0.f;
    r.scale = q.scale; r.hi = q.hi; r.alpha = q.alpha;
    r.blk0 = blocks;
    blocks += petch_blocks(q.nf, kBS);
    t.push_back(r);
  }
  if (t.empty()) return PETCH_OK;
  PETCH_ARG(ctx, blocks <= 0x7fffffffLL, "finish_round_batch: too many faces for one launch");
  const RoundItem* d =
      (const RoundItem*)petch_upload_table(ctx, t.data(), t.size() * sizeof(RoundItem));
  if (!d) return PETCH_E_HIP;
  hipLaunchKernelGGL(k_finish_round, dim3((unsigned)blocks), dim3(kBS), 0, ctx->stream, d,
                     (int)t.size());
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_finish_ion_batch(petch_ctx* ctx, int n, const petch_ion_finish_item* items) {
  PETCH_ARG(ctx, ctx && (n == 0 || items), "finish_ion_batch: null pointer");
  PETCH_ARG(ctx, n >= 0, "finish_ion_batch: n < 0");
  std::vector<IonItem> t;
  t.reserve((size_t)n);
  long long blocks = 0;
  for (int k = 0; k < n; ++k) {
    const petch_ion_finish_item& q = items[k];
    PETCH_ARG(ctx, q.nf >= 0 && q.acc && q.area && q.out, "finish_ion_batch: item %d is invalid",
              k);
    PETCH_ARG(ctx, !q.smooth || (q.nbr && q.normals), "finish_ion_batch: item %d needs the "
              "adjacency and normals to smooth", k);
    if (q.nf == 0) continue;
    IonItem r;
    r.acc = q.acc; r.area = q.area; r.nrm = q.normals; r.nbr = q.nbr; r.out = q.out;
    r.nf = q.nf;
    r.smooth = q.smooth && q.alpha > 0.f;
    r.scale = q.scale; r.alpha = q.alpha;
    r.blk0 = blocks;
    blocks += petch_blocks(q.nf, kBS);
    t.push_back(r);
  }
  if (t.empty()) return PETCH_OK;
  PETCH_ARG(ctx, blocks <= 0x7fffffffLL, "finish_ion_batch: too many faces for one launch");
  const IonItem* d = (const IonItem*)petch_upload_table(ctx, t.data(), t.size() * sizeof(IonItem));
  if (!d) return PETCH_E_HIP;
  hipLaunchKernelGGL(k_finish_ion, dim3((unsigned)blocks), dim3(kBS), 0, ctx->stream, d,
                     (int)t.size());
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_belen_coverage_adjoint(petch_ctx* ctx, const float* m_i, const float* m_F,
                                            const float* m_O, const float* cos_i,
                                            const float* iy_ie, const float* iy_ip,
                                            const petch_chem_params* cp, int F,
                                            const float* gbare, float* gm_F, float* gm_O,
                                            float* gm_i, float* g_ie, float* g_ip) {
  PETCH_ARG(ctx, ctx && m_F && m_O && cp && gbare && gm_F && gm_O,
            "belen_coverage_adjoint: null pointer");
  PETCH_ARG(ctx, (iy_ie && iy_ip) || (m_i && cos_i),
            "belen_coverage_adjoint: need ion yields or m_i/cos_i");
  if (F <= 0) return PETCH_OK;
  hipLaunchKernelGGL(k_belen_coverage_adjoint, dim3(petch_blocks(F, kBS)), dim3(kBS), 0,
                     ctx->stream, m_i, m_F, m_O, cos_i, iy_ie, iy_ip, make_chem(cp), F, gbare,
                     gm_F, gm_O, gm_i, g_ie, g_ip);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_surface_rate_adjoint(
    petch_ctx* ctx, const float* m_i, const float* m_F, const float* m_O, const float* cos_i,
    const float* iy_sp, const float* iy_ie, const float* iy_ip, const float* centroids,
    const uint8_t* is_mask, const petch_mask_geom* mg, const petch_chem_params* cp, int F,
    const float* gV, float* gm_F, float* gm_O, float* gm_i, float* g_sp, float* g_ie,
    float* g_ip, float* gY) {
  PETCH_ARG(ctx, ctx && m_i && m_F && m_O && cos_i && mg && cp && gV && gm_F && gm_O,
            "surface_rate_adjoint: null pointer");
  PETCH_ARG(ctx, centroids || is_mask, "surface_rate_adjoint: need centroids or an explicit mask");
  PETCH_ARG(ctx, !iy_ie || (iy_sp && iy_ip),
            "surface_rate_adjoint: ion yields need all three channels");
  if (F <= 0) return PETCH_OK;
  MaskGeom g{mg->sub_top, mg->sub_top + mg->mask_th, mg->Lx / 2.0, mg->Ly / 2.0,
             mg->trench_width / 2.0, (mg->trench_width / 2.0) * (mg->trench_width / 2.0),
             mg->hole};
  hipLaunchKernelGGL(k_surface_rate_adjoint, dim3(petch_blocks(F, kBS)), dim3(kBS), 0,
                     ctx->stream, m_i, m_F, m_O, cos_i, iy_sp, iy_ie, iy_ip, centroids, is_mask, g,
                     make_chem(cp), F, gV, gm_F, gm_O, gm_i, g_sp, g_ie, g_ip, gY);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_surface_rate(petch_ctx* ctx, const float* m_i, const float* m_F,
                                  const float* m_O, const float* cos_i, const float* iy_sp,
                                  const float* iy_ie, const float* iy_ip, const float* centroids,
                                  const uint8_t* is_mask, const petch_mask_geom* mg,
                                  const petch_chem_params* cp, int F, float* V, unsigned* red) {
  PETCH_ARG(ctx, ctx && m_i && m_F && m_O && cos_i && mg && cp && V && red,
            "surface_rate: null pointer");
  PETCH_ARG(ctx, centroids || is_mask, "surface_rate: need centroids or an explicit mask");
  PETCH_ARG(ctx, !iy_ie || (iy_sp && iy_ip), "surface_rate: ion yields need all three channels");
  PETCH_HIP(ctx, hipMemsetAsync(red, 0, 2 * sizeof(unsigned), ctx->stream));
  if (F <= 0) return PETCH_OK;
  MaskGeom g{mg->sub_top, mg->sub_top + mg->mask_th, mg->Lx / 2.0, mg->Ly / 2.0,
             mg->trench_width / 2.0, (mg->trench_width / 2.0) * (mg->trench_width / 2.0),
             mg->hole};
  hipLaunchKernelGGL(k_surface_rate, dim3(petch_blocks(F, kBS)), dim3(kBS), 0, ctx->stream, m_i,
                     m_F, m_O, cos_i, iy_sp, iy_ie, iy_ip, centroids, is_mask, g, make_chem(cp), F,
                     V, red);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_surface_rate_batch(petch_ctx* ctx, int n, const petch_rate_item* items) {
  PETCH_ARG(ctx, ctx && n >= 0 && (n == 0 || items), "surface_rate_batch: null pointer or n < 0");
  if (n == 0) return PETCH_OK;
  std::vector<RateItem> t((size_t)n);
  long long blocks = 0;
  for (int k = 0; k < n; ++k) {
    const petch_rate_item& a = items[k];
    PETCH_ARG(ctx, a.F >= 0 && a.red, "surface_rate_batch: item %d: F < 0 or no reduction pair", k);
    PETCH_ARG(ctx, a.F == 0 || (a.m_i && a.m_F && a.m_O && a.cos_i && a.V),
              "surface_rate_batch: item %d has a null pointer", k);
    PETCH_ARG(ctx, a.F == 0 || a.centroids || a.is_mask,
              "surface_rate_batch: item %d needs centroids or an explicit mask", k);
    PETCH_ARG(ctx, !a.iy_ie || (a.iy_sp && a.iy_ip),
              "surface_rate_batch: item %d: ion yields need all three channels", k);
    const petch_mask_geom& mg = a.mg;
    RateItem& r = t[k];
    r.m_i = a.m_i; r.m_F = a.m_F; r.m_O = a.m_O; r.cos_i = a.cos_i;
    r.iy_sp = a.iy_sp; r.iy_ie = a.iy_ie; r.iy_ip = a.iy_ip; r.centroids = a.centroids;
    r.is_mask = a.is_mask;
    r.mg = MaskGeom{mg.sub_top, mg.sub_top + mg.mask_th, mg.Lx / 2.0, mg.Ly / 2.0,
                    mg.trench_width / 2.0, (mg.trench_width / 2.0) * (mg.trench_width / 2.0),
                    mg.hole};
    r.c = make_chem(&a.chem);
    r.F = a.F;
    r.V = a.V;
    r.red = a.red;
    r.err_src = a.bvh ? petch_bvh_view(a.bvh).err : nullptr;
    r.err_dst = a.bvh_err;
    r.blk0 = blocks;
    blocks += petch_blocks(a.F, kBS);
  }
  PETCH_ARG(ctx, blocks <= 0x7fffffffLL, "surface_rate_batch: too many faces for one launch");
  const RateItem* d = (const RateItem*)petch_upload_table(ctx, t.data(), t.size() * sizeof(RateItem));
  if (!d) return PETCH_E_HIP;
  hipLaunchKernelGGL(k_rate_clear, dim3(petch_blocks(n, kBS)), dim3(kBS), 0, ctx->stream, d, n);
  PETCH_LAUNCH_CHECK(ctx);
  if (blocks > 0) {
    hipLaunchKernelGGL(k_rate_many, dim3((unsigned)blocks), dim3(kBS), 0, ctx->stream, d, n);
    PETCH_LAUNCH_CHECK(ctx);
  }
  return PETCH_OK;
}

// The band flags (zeroing Fs, which the extension then fills on the band), their scan and the
// band list of the items, then the extension queries; with gV_adjoint (one item) the
// transposed queries instead, which read the item's Fs as gFs (left untouched) and add it into
// gV_adjoint.
constexpr unsigned kExtendBlocks = 8192;

static int extend_run(petch_ctx* ctx, int n, const petch_extend_item* items, float* gV_adjoint) {
  PETCH_ARG(ctx, ctx && n >= 0 && (n == 0 || items), "extend_batch: null pointer or n < 0");
  if (n == 0) return PETCH_OK;
  std::vector<ExtendItem> t((size_t)n);
  long long total = 0;
  for (int f = 0; f < n; ++f) {
    const petch_extend_item& a = items[f];
    PETCH_ARG(ctx, a.bvh && a.V && a.phi && a.Fs, "extend_batch: item %d has a null pointer", f);
    PETCH_ARG(ctx, a.nx > 0 && a.ny > 0 && a.nz > 0, "extend_batch: item %d has a bad shape", f);
    total += (long long)a.nx * a.ny * a.nz;
    t[f] = ExtendItem{petch_bvh_view(a.bvh), a.V, a.phi, a.Fs, a.ny, a.nz,
                      petch::make_fastdiv((uint32_t)a.ny), petch::make_fastdiv((uint32_t)a.nz),
                      a.dx, a.band, total};
  }
  PETCH_ARG(ctx, total < 0x7fffffffLL, "extend_batch: %lld nodes exceed the int32 node index",
            total);
  const ExtendItem* dev = (const ExtendItem*)petch_upload_table(ctx, t.data(), t.size() * sizeof(ExtendItem));
  // the band's node list with its length after it, and the tile counts (scratch slots of the
  // level-set stages, which run after the extension on the same stream)
  const long long tiles = (total + kBandTile - 1) / kBandTile;
  int* idx = (int*)petch_scratch(ctx, SCR_LS_FROZEN, (size_t)(total + 1) * sizeof(int));
  int* tcnt = (int*)petch_scratch(ctx, SCR_LS_A, (size_t)tiles * sizeof(int));
  if (!dev || !idx || !tcnt) return PETCH_E_HIP;
  hipLaunchKernelGGL(k_band_count, dim3((unsigned)tiles), dim3(kBS), 0, ctx->stream, dev, n,
                     total, tcnt, gV_adjoint ? 0 : 1);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(petch::k_scan_partials, dim3(1), dim3(1024), 0, ctx->stream, tcnt,
                     (int)tiles, idx + total);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_band_emit, dim3((unsigned)tiles), dim3(kBS), 0, ctx->stream, dev, n, total,
                     tcnt, idx);
  PETCH_LAUNCH_CHECK(ctx);
  // the queries loop over the list (its length is on the device): enough workgroups to fill
  // every CU several times over, fewer when the whole grid has fewer nodes
  const unsigned g = petch_blocks(total, kBS);
  const unsigned gq = g < kExtendBlocks ? g : kExtendBlocks;
  if (gV_adjoint) {
    hipLaunchKernelGGL(k_extend_adjoint, dim3(gq), dim3(kBS), 0, ctx->stream, dev, n, idx,
                       idx + total, gV_adjoint);
  } else {
    hipLaunchKernelGGL(k_extend, dim3(gq), dim3(kBS), 0, ctx->stream, dev, n, idx, idx + total);
  }
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_extend_batch(petch_ctx* ctx, int n, const petch_extend_item* items) {
  return extend_run(ctx, n, items, nullptr);
}

extern "C" int petch_extend(petch_ctx* ctx, const petch_bvh* bvh, const float* V, const float* phi,
                            int nx, int ny, int nz, double dx, double band, float* Fs) {
  PETCH_ARG(ctx, ctx && bvh && V && phi && Fs, "extend: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "extend: bad shape");
  const petch_extend_item it{bvh, V, phi, Fs, nx, ny, nz, dx, band};
  return petch_extend_batch(ctx, 1, &it);
}

extern "C" int petch_extend_adjoint(petch_ctx* ctx, const petch_bvh* bvh, const float* phi, int nx,
                                    int ny, int nz, double dx, double band, const float* gFs,
                                    float* gV) {
  PETCH_ARG(ctx, ctx && bvh && phi && gFs && gV, "extend_adjoint: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "extend_adjoint: bad shape");
  PETCH_ARG(ctx, (const void*)gFs != (const void*)gV, "extend_adjoint: gFs and gV must differ");
  const long long N = (long long)nx * ny * nz;
  PETCH_ARG(ctx, N < 0x7fffffffLL, "extend_adjoint: %lld nodes exceed the int32 node index", N);
  PETCH_HIP(ctx, hipMemsetAsync(gV, 0, (size_t)bvh->nf * sizeof(float), ctx->stream));
  const petch_extend_item it{bvh, gV, phi, const_cast<float*>(gFs), nx, ny, nz, dx, band};
  return extend_run(ctx, 1, &it, gV);
}

extern "C" int petch_nearest_face_value_batch(petch_ctx* ctx, int n,
                                              const petch_nearest_item* items) {
  PETCH_ARG(ctx, ctx && n >= 0 && (n == 0 || items), "nearest_face_value_batch: null pointer or n < 0");
  std::vector<NearestItem> t;
  t.reserve((size_t)n);
  long long total = 0;
  for (int f = 0; f < n; ++f) {
    const petch_nearest_item& a = items[f];
    PETCH_ARG(ctx, a.n >= 0, "nearest_face_value_batch: item %d has n < 0", f);
    if (a.n == 0) continue;
    PETCH_ARG(ctx, a.bvh && a.src && a.pts && a.out,
              "nearest_face_value_batch: item %d has a null pointer", f);
    total += a.n;
    t.push_back(NearestItem{petch_bvh_view(a.bvh), a.src, a.pts, a.out, a.miss, total});
  }
  if (total == 0) return PETCH_OK;
  const NearestItem* dev =
      (const NearestItem*)petch_upload_table(ctx, t.data(), t.size() * sizeof(NearestItem));
  if (!dev) return PETCH_E_HIP;
  hipLaunchKernelGGL(k_nearest_batch, dim3(petch_blocks(total, kBS)), dim3(kBS), 0, ctx->stream,
                     dev, (int)t.size(), total);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_nearest_face_value(petch_ctx* ctx, const petch_bvh* bvh, const float* src,
                                        const float* pts, int n, float miss, float* out) {
  PETCH_ARG(ctx, ctx && bvh && src && pts && out, "nearest_face_value: null pointer");
  if (n <= 0) return PETCH_OK;
  const petch_nearest_item it{bvh, src, pts, n, miss, out};
  return petch_nearest_face_value_batch(ctx, 1, &it);
}
