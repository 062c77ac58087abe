// Per-face mesh attributes and edge adjacency.
//   petch_mesh_attrs    centroids (mean of the 3 vertices), areas (0.5 |cross|) and unit normals
//                       (threed.py:96-99, _face_normals_k threed.py:228-234)
//   petch_face_adjacency for each face and each of its edges (v0v1, v1v2, v2v0) the other face
//                       sharing that edge, or -1 (the pair set of _edge_adjacency,
//                       threed.py:103-125, as a gather-friendly 3-per-face table)
//   petch_gas_normals   face normals oriented into the gas by the sign of -grad(phi) at the
//                       centroid's nearest node (_gas_normals, threed.py:1553-1568)
#include "common.h"
#include "scan.h"

#include <vector>

namespace {

constexpr int kBS = 256;

// Writes face f's centroid, area and unit normal; returns whether the centroid is finite.
__device__ __forceinline__ bool mesh_attrs_face(const float* __restrict__ verts,
                                                const int* __restrict__ faces,
                                                float* __restrict__ cen, float* __restrict__ area,
                                                float* __restrict__ nrm, int f) {
  int a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
  float ax = verts[3 * a], ay = verts[3 * a + 1], az = verts[3 * a + 2];
  float bx = verts[3 * b], by = verts[3 * b + 1], bz = verts[3 * b + 2];
  float cx = verts[3 * c], cy = verts[3 * c + 1], cz = verts[3 * c + 2];
  const float gx = (ax + bx + cx) / 3.f, gy = (ay + by + cy) / 3.f, gz = (az + bz + cz) / 3.f;
  if (cen) {
    cen[3 * f + 0] = gx;
    cen[3 * f + 1] = gy;
    cen[3 * f + 2] = gz;
  }
  // cross product in float64: slivers (vertices near a grid node) keep an accurate normal
  double ux = (double)bx - ax, uy = (double)by - ay, uz = (double)bz - az;
  double vx = (double)cx - ax, vy = (double)cy - ay, vz = (double)cz - az;
  double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
  double len = sqrt(nx * nx + ny * ny + nz * nz);
  if (area) area[f] = (float)(0.5 * len);
  if (nrm) {
    double inv = len > 0.0 ? 1.0 / len : 0.0;
    nrm[3 * f + 0] = (float)(nx * inv);
    nrm[3 * f + 1] = (float)(ny * inv);
    nrm[3 * f + 2] = (float)(nz * inv);
  }
  return isfinite(gx) && isfinite(gy) && isfinite(gz);
}

__global__ void k_mesh_attrs(const float* __restrict__ verts, const int* __restrict__ faces, int nf,
                             float* __restrict__ cen, float* __restrict__ area,
                             float* __restrict__ nrm) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f < nf) mesh_attrs_face(verts, faces, cen, area, nrm, f);
}

__global__ void k_vert_count(const int* __restrict__ faces, int nf, int nv,
                             int* __restrict__ cnt) {
  int t = blockIdx.x * kBS + threadIdx.x;
  if (t >= 3 * nf) return;
  int v = faces[t];
  if (v >= 0 && v < nv) atomicAdd(&cnt[v], 1);
}

__device__ __forceinline__ void vert_fill_corner(const int* __restrict__ faces, int nv,
                                                 const int* __restrict__ off,
                                                 int* __restrict__ cur, int* __restrict__ inc,
                                                 int t) {
  int v = faces[t];
  if (v < 0 || v >= nv) return;
  int p = atomicAdd(&cur[v], 1);
  inc[off[v] + p] = t / 3;
}

__global__ void k_vert_fill(const int* __restrict__ faces, int nf, int nv,
                            const int* __restrict__ off, int* __restrict__ cur,
                            int* __restrict__ inc) {
  int t = blockIdx.x * kBS + threadIdx.x;
  if (t < 3 * nf) vert_fill_corner(faces, nv, off, cur, inc, t);
}

// Neighbour across edge (a,b) of face f: the smallest other face incident to a that also uses b.
__device__ __forceinline__ void face_nbr_edge(const int* __restrict__ faces, int nv,
                                              const int* __restrict__ off,
                                              const int* __restrict__ inc, int* __restrict__ nbr,
                                              int t) {
  int f = t / 3, e = t - 3 * f;
  int a = faces[3 * f + e], b = faces[3 * f + (e + 1) % 3];
  int best = -1;
  if (a >= 0 && a < nv) {
    for (int p = off[a], q = off[a + 1]; p < q; ++p) {
      int g = inc[p];
      if (g == f) continue;
      int g0 = faces[3 * g], g1 = faces[3 * g + 1], g2 = faces[3 * g + 2];
      if ((g0 == b || g1 == b || g2 == b) && (best < 0 || g < best)) best = g;
    }
  }
  nbr[t] = best;
}

__global__ void k_face_nbr(const int* __restrict__ faces, int nf, int nv, const int* __restrict__ off,
                           const int* __restrict__ inc, int* __restrict__ nbr) {
  int t = blockIdx.x * kBS + threadIdx.x;
  if (t < 3 * nf) face_nbr_edge(faces, nv, off, inc, nbr, t);
}

// The attributes and adjacency of many meshes (the features of a batched step), one launch per
// stage for all: item t owns the workgroups [blk0[t], blk0[t+1]) of the launches over its 3 nf
// face corners, and one workgroup per item scans its vertex counts. Every stage runs the device
// code of the single calls, so the outputs are theirs (the incidence lists fill in any order;
// the neighbour chosen is the smallest face id, whatever that order).
struct SurfItem {
  const float* verts;
  const int* faces;
  int nf, nv;
  float *cen, *area, *nrm;
  int* nbr;
  int *cnt, *cur, *off, *inc;   // workspace: nv + 1, nv + 1, nv + 1, 3 nf
  int* finite;                  // optional: 1 if every centroid is finite, else 0
  long long blk0;
};

__device__ __forceinline__ int surf_item_of(const SurfItem* __restrict__ it, int n) {
  int lo = 0, hi = n - 1;
  const long long b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (it[mid].blk0 <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// one workgroup per item: clear the vertex counters and set the finite flag
__global__ void k_surf_clear(const SurfItem* __restrict__ items) {
  const SurfItem& it = items[blockIdx.x];
  if (threadIdx.x == 0 && it.finite) *it.finite = 1;
  for (int v = threadIdx.x; v <= it.nv; v += kBS) {
    it.cnt[v] = 0;
    it.cur[v] = 0;
  }
}

// corner t: its vertex's count, and (first corner of a face) the face's attributes
__global__ void k_surf_count(const SurfItem* __restrict__ items, int n) {
  const SurfItem& it = items[surf_item_of(items, n)];
  const int t = (int)(blockIdx.x - it.blk0) * kBS + threadIdx.x;
  if (t >= 3 * it.nf) return;
  const int v = it.faces[t];
  if (v >= 0 && v < it.nv) atomicAdd(&it.cnt[v], 1);
  if (t % 3 == 0 && !mesh_attrs_face(it.verts, it.faces, it.cen, it.area, it.nrm, t / 3) &&
      it.finite)
    *it.finite = 0;
}

// one workgroup per item: exclusive scan of its nv + 1 counts
__global__ void k_surf_scan(const SurfItem* __restrict__ items) {
  const SurfItem& it = items[blockIdx.x];
  const int m = it.nv + 1;
  int carry = 0;
  for (int c = 0; c < m; c += kBS) {
    const int i = c + threadIdx.x;
    const int v = i < m ? it.cnt[i] : 0;
    int tot;
    const int ex = petch::block_excl_scan<kBS>(v, tot);
    if (i < m) it.off[i] = carry + ex;
    carry += tot;
  }
}

__global__ void k_surf_fill(const SurfItem* __restrict__ items, int n) {
  const SurfItem& it = items[surf_item_of(items, n)];
  const int t = (int)(blockIdx.x - it.blk0) * kBS + threadIdx.x;
  if (t < 3 * it.nf) vert_fill_corner(it.faces, it.nv, it.off, it.cur, it.inc, t);
}

__global__ void k_surf_nbr(const SurfItem* __restrict__ items, int n) {
  const SurfItem& it = items[surf_item_of(items, n)];
  const int t = (int)(blockIdx.x - it.blk0) * kBS + threadIdx.x;
  if (t < 3 * it.nf) face_nbr_edge(it.faces, it.nv, it.off, it.inc, it.nbr, t);
}

__global__ void k_gas_normals(const float* __restrict__ phi, int nx, int ny, int nz, float dx,
                              const float* __restrict__ cen, const float* __restrict__ nrm_in,
                              int nf, float* __restrict__ out) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f >= nf) return;
  float inv = 1.f / dx;
  int idx[3], dim[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    int q = (int)rintf(cen[3 * f + a] * inv);
    idx[a] = q < 0 ? 0 : (q > dim[a] - 1 ? dim[a] - 1 : q);
  }
  long long sx = (long long)ny * nz, sy = nz;
  long long n = idx[0] * sx + idx[1] * sy + idx[2];
  long long st[3] = {sx, sy, 1};
  float g[3];
  // numpy.gradient: central inside, one-sided first order at the two ends
  for (int a = 0; a < 3; ++a) {
    if (dim[a] < 2) {
      g[a] = 0.f;
    } else if (idx[a] == 0) {
      g[a] = (phi[n + st[a]] - phi[n]) * inv;
    } else if (idx[a] == dim[a] - 1) {
      g[a] = (phi[n] - phi[n - st[a]]) * inv;
    } else {
      g[a] = (phi[n + st[a]] - phi[n - st[a]]) * 0.5f * inv;
    }
  }
  float nx_ = nrm_in[3 * f], ny_ = nrm_in[3 * f + 1], nz_ = nrm_in[3 * f + 2];
  float d = -(nx_ * g[0] + ny_ * g[1] + nz_ * g[2]);
  float s = d < 0.f ? -1.f : 1.f;
  out[3 * f] = s * nx_;
  out[3 * f + 1] = s * ny_;
  out[3 * f + 2] = s * nz_;
}

}  // namespace

extern "C" int petch_mesh_attrs(petch_ctx* ctx, const float* verts, const int* faces, int nf,
                                float* centroids, float* areas, float* normals) {
  PETCH_ARG(ctx, ctx && verts && faces, "mesh_attrs: null pointer");
  PETCH_ARG(ctx, nf >= 0, "mesh_attrs: nf < 0");
  if (nf == 0) return PETCH_OK;
  hipLaunchKernelGGL(k_mesh_attrs, dim3(petch_blocks(nf, kBS)), dim3(kBS), 0, ctx->stream, verts,
                     faces, nf, centroids, areas, normals);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_face_adjacency(petch_ctx* ctx, const int* faces, int nf, int nv, int* nbr) {
  PETCH_ARG(ctx, ctx && faces && nbr, "face_adjacency: null pointer");
  PETCH_ARG(ctx, nf >= 0 && nv >= 0, "face_adjacency: negative size");
  if (nf == 0) return PETCH_OK;
  // layout: cnt/cur [nv+1] | off [nv+1] | total | inc [3nf]
  size_t n1 = (size_t)nv + 1;
  int* w = (int*)petch_scratch(ctx, SCR_ADJ, (2 * n1 + 1 + 3 * (size_t)nf) * sizeof(int));
  if (!w) return PETCH_E_HIP;
  int* cnt = w;
  int* off = w + n1;
  int* total = w + 2 * n1;
  int* inc = w + 2 * n1 + 1;
  PETCH_HIP(ctx, hipMemsetAsync(cnt, 0, n1 * sizeof(int), ctx->stream));
  unsigned g3 = petch_blocks(3LL * nf, kBS);
  hipLaunchKernelGGL(k_vert_count, dim3(g3), dim3(kBS), 0, ctx->stream, faces, nf, nv, cnt);
  PETCH_LAUNCH_CHECK(ctx);
  int rc = petch_exclusive_scan(ctx, cnt, off, (long long)n1, total);
  if (rc != PETCH_OK) return rc;
  PETCH_HIP(ctx, hipMemsetAsync(cnt, 0, n1 * sizeof(int), ctx->stream));
  hipLaunchKernelGGL(k_vert_fill, dim3(g3), dim3(kBS), 0, ctx->stream, faces, nf, nv, off, cnt,
                     inc);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_face_nbr, dim3(g3), dim3(kBS), 0, ctx->stream, faces, nf, nv, off, inc,
                     nbr);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_surface_build_batch(petch_ctx* ctx, int n, const petch_surface_item* items) {
  PETCH_ARG(ctx, ctx && (n == 0 || items), "surface_build_batch: null pointer");
  PETCH_ARG(ctx, n >= 0, "surface_build_batch: n < 0");
  std::vector<SurfItem> t;
  t.reserve((size_t)n);
  long long ws = 0, blocks = 0;
  for (int k = 0; k < n; ++k) {
    const petch_surface_item& q = items[k];
    PETCH_ARG(ctx, q.nf >= 0 && q.nv >= 0, "surface_build_batch: item %d has a negative size", k);
    if (q.nf == 0) continue;
    PETCH_ARG(ctx, q.verts && q.faces && q.centroids && q.areas && q.normals && q.nbr,
              "surface_build_batch: item %d has a null pointer", k);
    ws += 3LL * (q.nv + 1) + 3LL * q.nf;
  }
  int* w = ws ? (int*)petch_scratch(ctx, SCR_ADJ, (size_t)ws * sizeof(int)) : nullptr;
  if (ws && !w) return PETCH_E_HIP;
  long long off = 0;
  for (int k = 0; k < n; ++k) {
    const petch_surface_item& q = items[k];
    if (q.nf == 0) continue;
    SurfItem r;
    r.verts = q.verts; r.faces = q.faces; r.nf = q.nf; r.nv = q.nv;
    r.cen = q.centroids; r.area = q.areas; r.nrm = q.normals; r.nbr = q.nbr;
    r.cnt = w + off; off += q.nv + 1;
    r.cur = w + off; off += q.nv + 1;
    r.off = w + off; off += q.nv + 1;
    r.inc = w + off; off += 3LL * q.nf;
    r.finite = q.finite;
    r.blk0 = blocks;
    blocks += petch_blocks(3LL * q.nf, kBS);
    t.push_back(r);
  }
  if (t.empty()) return PETCH_OK;
  PETCH_ARG(ctx, blocks <= 0x7fffffffLL, "surface_build_batch: too many faces for one launch");
  const SurfItem* d = (const SurfItem*)petch_upload_table(ctx, t.data(), t.size() * sizeof(SurfItem));
  if (!d) return PETCH_E_HIP;
  const int m = (int)t.size();
  hipLaunchKernelGGL(k_surf_clear, dim3((unsigned)m), dim3(kBS), 0, ctx->stream, d);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_surf_count, dim3((unsigned)blocks), dim3(kBS), 0, ctx->stream, d, m);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_surf_scan, dim3((unsigned)m), dim3(kBS), 0, ctx->stream, d);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_surf_fill, dim3((unsigned)blocks), dim3(kBS), 0, ctx->stream, d, m);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_surf_nbr, dim3((unsigned)blocks), dim3(kBS), 0, ctx->stream, d, m);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_gas_normals(petch_ctx* ctx, const float* phi, int nx, int ny, int nz,
                                 double dx, const float* centroids, const float* normals, int nf,
                                 float* out) {
  PETCH_ARG(ctx, ctx && phi && centroids && normals && out, "gas_normals: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0 && nf >= 0, "gas_normals: bad size");
  if (nf == 0) return PETCH_OK;
  hipLaunchKernelGGL(k_gas_normals, dim3(petch_blocks(nf, kBS)), dim3(kBS), 0, ctx->stream, phi,
                     nx, ny, nz, (float)dx, centroids, normals, nf, out);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}
