// Morton-code LBVH over the surface triangles, rebuilt every step (replaces
// wp.Mesh(points, indices), threed.py:1674-1675):
//   prep     triangle centroids + per-block bounds        -> scene bounds (one workgroup)
//   morton   30-bit codes of the centroids                -> stable LDS radix sort (sort.hip)
//   leaves   triangle vertices gathered into Morton order (48 B per leaf)
//   karras   internal nodes from the sorted codes (Karras, HPG 2012), ties broken by index
//   refit    leaves climb to the root; at each node the second arriving thread (atomic flag)
//            unions the two child boxes and continues, so each node is built exactly once
//   wide     4-wide nodes at binary-node indices: on large meshes a top-down collapse that
//            opens the larger-area inner child first (one workgroup, level by level); on small
//            ones every binary node at even depth (parity from its parent chain) gathers its
//            grandchildren (one parallel pass)
// plus the two query entry points used by the tests (the trace kernels inline the traversal).
#include "common.h"
#include <vector>

#include <stdlib.h>
#include "bvh.h"
#include "scan.h"

int petch_radix_sort_pairs(petch_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* ktmp,
                           uint32_t* vtmp, int n, int bits);

using petch::BvhNode;
using petch::BvhView;
using petch::BvhWide;

namespace {

constexpr int kBS = 256;

struct Work {  // device scratch owned by the BVH, carved from bvh->work
  float4* cen;
  uint32_t *keys, *vals, *ktmp, *vtmp;
  int *pint, *pleaf, *flags, *err;
  float* part;   // per-block bounds (6 per block)
  float* scene;  // 6 floats
  float* root;   // 6 floats
};

size_t work_bytes(int cap_f) {
  size_t nb = (cap_f + kBS - 1) / kBS;
  return (size_t)cap_f * 16 + (size_t)cap_f * 4 * 4 + (size_t)cap_f * 4 * 3 + 64 +
         nb * 6 * 4 + 64 + 64 + 1024;
}

Work carve(void* base, int cap_f) {
  char* p = (char*)base;
  Work w;
  size_t nb = (cap_f + kBS - 1) / kBS;
  w.cen = (float4*)p; p += (size_t)cap_f * 16;
  w.keys = (uint32_t*)p; p += (size_t)cap_f * 4;
  w.vals = (uint32_t*)p; p += (size_t)cap_f * 4;
  w.ktmp = (uint32_t*)p; p += (size_t)cap_f * 4;
  w.vtmp = (uint32_t*)p; p += (size_t)cap_f * 4;
  w.pint = (int*)p; p += (size_t)cap_f * 4;
  w.pleaf = (int*)p; p += (size_t)cap_f * 4;
  w.flags = (int*)p; p += (size_t)cap_f * 4;
  w.err = (int*)p; p += 64;
  w.part = (float*)p; p += nb * 6 * 4;
  w.scene = (float*)p; p += 64;
  w.root = (float*)p; p += 64;
  return w;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
  return v;
}

// Centroids of block lb's faces and the block's centroid bounds (6 floats at part[6 lb]).
__device__ __forceinline__ void bvh_prep_block(const float* __restrict__ verts,
                                               const int* __restrict__ faces, int nf,
                                               float4* __restrict__ cen, float* __restrict__ part,
                                               int lb) {
  int f = lb * kBS + threadIdx.x;
  float c[3] = {3.0e38f, 3.0e38f, 3.0e38f}, cmax[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
  if (f < nf) {
    int a = faces[3 * f], b = faces[3 * f + 1], d = faces[3 * f + 2];
    float x = (verts[3 * a] + verts[3 * b] + verts[3 * d]) / 3.f;
    float y = (verts[3 * a + 1] + verts[3 * b + 1] + verts[3 * d + 1]) / 3.f;
    float z = (verts[3 * a + 2] + verts[3 * b + 2] + verts[3 * d + 2]) / 3.f;
    cen[f] = make_float4(x, y, z, 0.f);
    c[0] = cmax[0] = x;
    c[1] = cmax[1] = y;
    c[2] = cmax[2] = z;
  }
  __shared__ float s[kBS / 64][6];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < 3; ++k) {
    float mn = wave_min(c[k]), mx = wave_max(cmax[k]);
    if (lane == 0) {
      s[w][k] = mn;
      s[w][3 + k] = mx;
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    int k = threadIdx.x;
    float v = s[0][k];
    for (int q = 1; q < kBS / 64; ++q) v = (k < 3) ? fminf(v, s[q][k]) : fmaxf(v, s[q][k]);
    part[lb * 6 + k] = v;
  }
}

__global__ void __launch_bounds__(kBS) k_bvh_prep(const float* __restrict__ verts,
                                                  const int* __restrict__ faces, int nf,
                                                  float4* __restrict__ cen,
                                                  float* __restrict__ part) {
  bvh_prep_block(verts, faces, nf, cen, part, blockIdx.x);
}

// The scene bounds from nb block partials (one workgroup).
__device__ __forceinline__ void bvh_bounds_wg(const float* __restrict__ part, int nb,
                                              float* __restrict__ scene) {
  __shared__ float s[6][kBS];
  for (int k = 0; k < 6; ++k) {
    float v = (k < 3) ? 3.0e38f : -3.0e38f;
    for (int b = threadIdx.x; b < nb; b += kBS)
      v = (k < 3) ? fminf(v, part[b * 6 + k]) : fmaxf(v, part[b * 6 + k]);
    s[k][threadIdx.x] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    int k = threadIdx.x;
    float v = s[k][0];
    for (int q = 1; q < kBS; ++q) v = (k < 3) ? fminf(v, s[k][q]) : fmaxf(v, s[k][q]);
    scene[k] = v;
  }
}

__global__ void __launch_bounds__(kBS) k_bvh_bounds(const float* __restrict__ part, int nb,
                                                    float* __restrict__ scene) {
  bvh_bounds_wg(part, nb, scene);
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__device__ __forceinline__ void bvh_morton_face(const float4* __restrict__ cen,
                                                const float* __restrict__ scene,
                                                uint32_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals, int f) {
  float4 c = cen[f];
  float q[3] = {c.x, c.y, c.z};
  uint32_t u[3];
  for (int k = 0; k < 3; ++k) {
    float ext = scene[3 + k] - scene[k];
    float t = ext > 0.f ? (q[k] - scene[k]) / ext : 0.f;
    float s = fminf(fmaxf(t * 1024.f, 0.f), 1023.f);
    u[k] = (uint32_t)s;
  }
  keys[f] = (spread10(u[0]) << 2) | (spread10(u[1]) << 1) | spread10(u[2]);
  vals[f] = (uint32_t)f;
}

__global__ void k_bvh_morton(const float4* __restrict__ cen, int nf,
                             const float* __restrict__ scene, uint32_t* __restrict__ keys,
                             uint32_t* __restrict__ vals) {
  int f = blockIdx.x * kBS + threadIdx.x;
  if (f < nf) bvh_morton_face(cen, scene, keys, vals, f);
}

__device__ __forceinline__ void bvh_leaf(const float* __restrict__ verts,
                                         const int* __restrict__ faces,
                                         const uint32_t* __restrict__ vals,
                                         float4* __restrict__ tris, int s) {
  int f = (int)vals[s];
  int a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
  tris[3 * s] = make_float4(verts[3 * a], verts[3 * a + 1], verts[3 * a + 2], __int_as_float(f));
  tris[3 * s + 1] = make_float4(verts[3 * b], verts[3 * b + 1], verts[3 * b + 2], 0.f);
  tris[3 * s + 2] = make_float4(verts[3 * c], verts[3 * c + 1], verts[3 * c + 2], 0.f);
}

__global__ void k_bvh_leaves(const float* __restrict__ verts, const int* __restrict__ faces,
                             const uint32_t* __restrict__ vals, int nf,
                             float4* __restrict__ tris) {
  int s = blockIdx.x * kBS + threadIdx.x;
  if (s < nf) bvh_leaf(verts, faces, vals, tris, s);
}

__device__ __forceinline__ int delta(const uint32_t* __restrict__ k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = k[i], b = k[j];
  if (a == b) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(a ^ b);
}

__device__ __forceinline__ void bvh_karras_node(const uint32_t* __restrict__ keys, int n,
                                                BvhNode* __restrict__ nodes, int* __restrict__ pint,
                                                int* __restrict__ pleaf, int i) {
  int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta(keys, n, i, j);
  int s = 0, t = l;
  while (true) {
    t = (t + 1) >> 1;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  int gamma = i + s * d + (d < 0 ? -1 : 0);
  int lo = i < j ? i : j, hi = i < j ? j : i;
  int left = (lo == gamma) ? ~gamma : gamma;
  int right = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
  nodes[i].d = make_int4(left, right, 0, 0);
  if (left < 0) pleaf[~left] = i; else pint[left] = i;
  if (right < 0) pleaf[~right] = i; else pint[right] = i;
  if (i == 0) pint[0] = -1;
}

__global__ void k_bvh_karras(const uint32_t* __restrict__ keys, int n, BvhNode* __restrict__ nodes,
                             int* __restrict__ pint, int* __restrict__ pleaf) {
  int i = blockIdx.x * kBS + threadIdx.x;
  if (i < n - 1) bvh_karras_node(keys, n, nodes, pint, pleaf, i);
}

__device__ __forceinline__ void put_child_box(BvhNode* nd, bool left, const float lo[3],
                                              const float hi[3]) {
  float* p = (float*)nd;
  if (left) {
    p[0] = lo[0]; p[1] = lo[1]; p[2] = lo[2]; p[3] = hi[0]; p[4] = hi[1]; p[5] = hi[2];
  } else {
    p[6] = lo[0]; p[7] = lo[1]; p[8] = lo[2]; p[9] = hi[0]; p[10] = hi[1]; p[11] = hi[2];
  }
}

// Leaf s climbs towards the root; at each node the second arriving thread (acq_rel flag at
// memory scope kScope) writes the node's box from both children and goes on, the first one
// stops. Threads of the whole grid need agent scope; within one workgroup (all threads on one
// CU, one L2) workgroup scope suffices and spares the L2 write-back an agent-scope release
// costs on a chip whose XCDs keep separate L2s.
template <int kScope = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ void refit_leaf(const float4* __restrict__ tris, BvhNode* nodes,
                                           const int* __restrict__ pint,
                                           const int* __restrict__ pleaf, int* flags, int n,
                                           float* root, int s) {
  float4 a = tris[3 * s], b = tris[3 * s + 1], c = tris[3 * s + 2];
  float lo[3] = {fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z))};
  float hi[3] = {fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z))};
  int child = ~s;
  int p = pleaf[s];
  for (int guard = 0; p >= 0 && p < n - 1 && guard < n; ++guard) {
    BvhNode* nd = nodes + p;
    bool isleft = nd->d.x == child;
    put_child_box(nd, isleft, lo, hi);
    int old = __hip_atomic_fetch_add(flags + p, 1, __ATOMIC_ACQ_REL, kScope);
    if (old == 0) return;  // the sibling's thread finishes this node
    const float* q = (const float*)nd;
    for (int k = 0; k < 3; ++k) {
      lo[k] = fminf(q[k], q[6 + k]);
      hi[k] = fmaxf(q[3 + k], q[9 + k]);
    }
    child = p;
    p = pint[p];
  }
  for (int k = 0; k < 3; ++k) {
    root[k] = lo[k];
    root[3 + k] = hi[k];
  }
}

__global__ void k_bvh_refit(const float4* __restrict__ tris, BvhNode* nodes,
                            const int* __restrict__ pint, const int* __restrict__ pleaf,
                            int* flags, int n, float* root) {
  int s = blockIdx.x * kBS + threadIdx.x;
  if (s >= n) return;
  refit_leaf(tris, nodes, pint, pleaf, flags, n, root, s);
}

// The refits of many BVHs in one launch: the climbs are chains of dependent global accesses
// (tens of microseconds for one tree however small), so the trees of a batched step refit side
// by side instead of one after the other.
struct RefitItem {
  const float4* tris;
  BvhNode* nodes;
  const int *pint, *pleaf;
  int* flags;
  float* root;
  int n;
};

// One workgroup per tree (the trees of a batched build): the workgroup's threads climb from
// all of the tree's leaves in turn with workgroup-scope flags, and the trees refit side by side.
constexpr int kRefitWG = 1024;
__global__ void __launch_bounds__(kRefitWG) k_bvh_refit_wg(const RefitItem* __restrict__ items) {
  const RefitItem& it = items[blockIdx.x];
  for (int s = threadIdx.x; s < it.n; s += kRefitWG)
    refit_leaf<__HIP_MEMORY_SCOPE_WORKGROUP>(it.tris, it.nodes, it.pint, it.pleaf, it.flags, it.n,
                                             it.root, s);
}

__device__ __forceinline__ void bvh_wide_node(const BvhNode* __restrict__ nodes,
                                              const int* __restrict__ pint, int n,
                                              BvhWide* __restrict__ wide, int i) {
  int depth = 0;
  int p = pint[i];
  for (int guard = 0; p >= 0 && p < n - 1 && guard < n; ++guard) {
    ++depth;
    p = pint[p];
  }
  if (depth & 1) return;
  float lo[4][3], hi[4][3];
  int ref[4];
  int m = 0;
  const BvhNode& N = nodes[i];
  const int kids[2] = {N.d.x, N.d.y};
  for (int side = 0; side < 2; ++side) {
    const int c = kids[side];
    if (c < 0 || c >= n - 1) {  // a leaf (an out-of-range reference is kept for the guard)
      const float* q = (const float*)&N + 6 * side;
      for (int k = 0; k < 3; ++k) {
        lo[m][k] = q[k];
        hi[m][k] = q[3 + k];
      }
      ref[m++] = c;
    } else {
      const BvhNode& C = nodes[c];
      const float* q = (const float*)&C;
      for (int g = 0; g < 2; ++g) {
        for (int k = 0; k < 3; ++k) {
          lo[m][k] = q[6 * g + k];
          hi[m][k] = q[6 * g + 3 + k];
        }
        ref[m++] = g == 0 ? C.d.x : C.d.y;
      }
    }
  }
  for (; m < 4; ++m) {
    for (int k = 0; k < 3; ++k) {
      lo[m][k] = INFINITY;
      hi[m][k] = INFINITY;
    }
    ref[m] = petch::kWideEmpty;
  }
  BvhWide w;
  w.lox = make_float4(lo[0][0], lo[1][0], lo[2][0], lo[3][0]);
  w.hix = make_float4(hi[0][0], hi[1][0], hi[2][0], hi[3][0]);
  w.loy = make_float4(lo[0][1], lo[1][1], lo[2][1], lo[3][1]);
  w.hiy = make_float4(hi[0][1], hi[1][1], hi[2][1], hi[3][1]);
  w.loz = make_float4(lo[0][2], lo[1][2], lo[2][2], lo[3][2]);
  w.hiz = make_float4(hi[0][2], hi[1][2], hi[2][2], hi[3][2]);
  w.ref = make_int4(ref[0], ref[1], ref[2], ref[3]);
  w.pad = make_int4(0, 0, 0, 0);
  wide[i] = w;
}

__global__ void k_bvh_wide(const BvhNode* __restrict__ nodes, const int* __restrict__ pint, int n,
                           BvhWide* __restrict__ wide) {
  int i = blockIdx.x * kBS + threadIdx.x;
  if (i < n - 1) bvh_wide_node(nodes, pint, n, wide, i);
}

// The Morton-tree builds of many meshes (the features of a batched step) side by side: one
// launch per stage for all of them. Item t owns the workgroups [fblk0[t], fblk0[t+1]) of the
// launches over faces and [nblk0[t], nblk0[t+1]) of those over internal nodes; every stage runs
// the same device code as the single build, so the trees are the same.
struct BuildItem {
  const float* verts;
  const int* faces;
  int nf, nfb;               // faces, and workgroups over them
  Work w;
  BvhNode* nodes;
  BvhWide* wide;
  float4* tris;
  long long fblk0, nblk0;
};

__device__ __forceinline__ int build_item_of(const BuildItem* __restrict__ it, int n, bool nodes) {
  int lo = 0, hi = n - 1;
  const long long b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((nodes ? it[mid].nblk0 : it[mid].fblk0) <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(kBS) k_bvh_prep_many(const BuildItem* __restrict__ items, int n) {
  const BuildItem& it = items[build_item_of(items, n, false)];
  bvh_prep_block(it.verts, it.faces, it.nf, it.w.cen, it.w.part, (int)(blockIdx.x - it.fblk0));
}

// one workgroup per item: the scene bounds, and the traversal error flag cleared
__global__ void __launch_bounds__(kBS) k_bvh_bounds_many(const BuildItem* __restrict__ items) {
  const BuildItem& it = items[blockIdx.x];
  bvh_bounds_wg(it.w.part, it.nfb, it.w.scene);
  if (threadIdx.x == 0) *it.w.err = 0;
}

__global__ void k_bvh_morton_many(const BuildItem* __restrict__ items, int n) {
  const BuildItem& it = items[build_item_of(items, n, false)];
  const int f = (int)(blockIdx.x - it.fblk0) * kBS + threadIdx.x;
  if (f < it.nf) bvh_morton_face(it.w.cen, it.w.scene, it.w.keys, it.w.vals, f);
}

// the leaves in Morton order, and the refit flags cleared
__global__ void k_bvh_leaves_many(const BuildItem* __restrict__ items, int n) {
  const BuildItem& it = items[build_item_of(items, n, false)];
  const int f = (int)(blockIdx.x - it.fblk0) * kBS + threadIdx.x;
  if (f >= it.nf) return;
  bvh_leaf(it.verts, it.faces, it.w.vals, it.tris, f);
  it.w.flags[f] = 0;
}

__global__ void k_bvh_karras_many(const BuildItem* __restrict__ items, int n) {
  const BuildItem& it = items[build_item_of(items, n, true)];
  const int i = (int)(blockIdx.x - it.nblk0) * kBS + threadIdx.x;
  if (i < it.nf - 1) bvh_karras_node(it.w.keys, it.nf, it.nodes, it.w.pint, it.w.pleaf, i);
}

__global__ void k_bvh_wide_many(const BuildItem* __restrict__ items, int n) {
  const BuildItem& it = items[build_item_of(items, n, true)];
  const int i = (int)(blockIdx.x - it.nblk0) * kBS + threadIdx.x;
  if (i < it.nf - 1) bvh_wide_node(it.nodes, it.w.pint, it.nf, it.wide, i);
}

// Surface area of the box lo..hi (0 for an inverted box).
__device__ __forceinline__ float box_area(const float* lo, const float* hi) {
  const float x = fmaxf(hi[0] - lo[0], 0.f), y = fmaxf(hi[1] - lo[1], 0.f),
              z = fmaxf(hi[2] - lo[2], 0.f);
  return x * y + y * z + z * x;
}

constexpr int kWideBS = 1024;
constexpr int kWideMaxLevels = 96;
// The area-driven collapse runs in one workgroup (~0.28 ms at 47k faces) and the clustering in
// ~50 launches (~1 ms); together they save ~7 % of trace time. Unless the caller chose a
// quality (petch_bvh_set_quality), meshes below this size get the Morton tree and the
// per-level grandchildren rule (one parallel pass).
constexpr int kWideSahMinFaces = 32768;

// Top-down collapse into 4-wide nodes, one workgroup: starting from the root, each wide node
// takes its binary node's two children and, twice, replaces the inner child of largest surface
// area by that child's two children (a leaf stays a leaf), so big boxes are split first. The
// wide node is stored at its binary node's index and its inner children form the next level.
// A tree deeper than kWideMaxLevels wide levels leaves error flag 3 (never for Morton trees,
// whose depth is bounded by the code length plus the index tie-break).
__global__ void __launch_bounds__(kWideBS) k_bvh_wide_sah(const BvhNode* __restrict__ nodes,
                                                          int n, BvhWide* __restrict__ wide,
                                                          int* __restrict__ list_a,
                                                          int* __restrict__ list_b,
                                                          int* __restrict__ err) {
  __shared__ int cnt_cur, cnt_next;
  if (threadIdx.x == 0) {
    list_a[0] = 0;
    cnt_cur = 1;
    cnt_next = 0;
  }
  __syncthreads();
  int* cur = list_a;
  int* nxt = list_b;
  for (int level = 0; level < kWideMaxLevels; ++level) {
    const int m = cnt_cur;
    if (m == 0) break;
    for (int t = threadIdx.x; t < m; t += kWideBS) {
      const int b = cur[t];
      float lo[4][3], hi[4][3];
      int ref[4];
      const float* q = (const float*)(nodes + b);
      for (int side = 0; side < 2; ++side) {
        for (int k = 0; k < 3; ++k) {
          lo[side][k] = q[6 * side + k];
          hi[side][k] = q[6 * side + 3 + k];
        }
        ref[side] = side == 0 ? nodes[b].d.x : nodes[b].d.y;
      }
      int cnt = 2;
      while (cnt < 4) {
        int best = -1;
        float best_a = -1.f;
        for (int c = 0; c < cnt; ++c) {
          if (ref[c] < 0 || ref[c] >= n - 1) continue;   // leaves (and bad refs) stay
          const float a = box_area(lo[c], hi[c]);
          if (a > best_a) {
            best_a = a;
            best = c;
          }
        }
        if (best < 0) break;
        const int r = ref[best];
        const float* g = (const float*)(nodes + r);
        for (int k = 0; k < 3; ++k) {
          lo[best][k] = g[k];
          hi[best][k] = g[3 + k];
          lo[cnt][k] = g[6 + k];
          hi[cnt][k] = g[9 + k];
        }
        ref[best] = nodes[r].d.x;
        ref[cnt] = nodes[r].d.y;
        ++cnt;
      }
      for (int c = cnt; c < 4; ++c) {
        for (int k = 0; k < 3; ++k) {
          lo[c][k] = INFINITY;
          hi[c][k] = INFINITY;
        }
        ref[c] = petch::kWideEmpty;
      }
      BvhWide w;
      w.lox = make_float4(lo[0][0], lo[1][0], lo[2][0], lo[3][0]);
      w.hix = make_float4(hi[0][0], hi[1][0], hi[2][0], hi[3][0]);
      w.loy = make_float4(lo[0][1], lo[1][1], lo[2][1], lo[3][1]);
      w.hiy = make_float4(hi[0][1], hi[1][1], hi[2][1], hi[3][1]);
      w.loz = make_float4(lo[0][2], lo[1][2], lo[2][2], lo[3][2]);
      w.hiz = make_float4(hi[0][2], hi[1][2], hi[2][2], hi[3][2]);
      w.ref = make_int4(ref[0], ref[1], ref[2], ref[3]);
      w.pad = make_int4(0, 0, 0, 0);
      wide[b] = w;
      for (int c = 0; c < cnt; ++c) {
        if (ref[c] >= 0 && ref[c] < n - 1) {
          const int slot = atomicAdd(&cnt_next, 1);
          if (slot < n - 1) nxt[slot] = ref[c];
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      cnt_cur = cnt_next < n - 1 ? cnt_next : n - 1;
      cnt_next = 0;
    }
    __syncthreads();
    int* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  if (threadIdx.x == 0 && cnt_cur > 0 && err) *err = 3;
}

// ---------------------------------------------------------------------------- PLOC
// Parallel locally-ordered clustering (Meister & Bittner, IEEE TVCG 2018) over the Morton-sorted
// leaves: every cluster finds, within kPlocRadius places of it in Morton order, the cluster whose
// union box has the smallest surface area; mutual nearest pairs merge into a new internal node;
// the survivors are compacted in order, and the rounds repeat until one cluster is left. Node
// indices count down from n-2, so the last merge is the root, node 0, as in the Morton tree.
// Builds a tree closer to a surface-area-optimal one than the Morton split. The first rounds
// run as grid-wide launches; the remaining few thousand clusters are finished by one workgroup
// in a single launch. If the rounds do not get down to one cluster the Morton tree is kept.
constexpr int kPlocRadius = 16;
constexpr int kPlocRounds = 6;          // rounds as grid-wide launches (the big early rounds)
constexpr int kPlocFinishRounds = 512;  // then one workgroup finishes the (small) rest

struct PlocWork {
  int *ca, *cb;     // cluster refs (>= 0 node, < 0 leaf ~slot), ping-pong
  float4 *ba, *bb;  // cluster boxes (lo, hi) as two float4 per cluster, ping-pong
  int *nn, *valid, *pos;
  int* cnt;         // cluster counts, one slot per round parity
  int* alloc;       // merges so far
  BvhNode* nodes;   // the clustered tree
};

size_t ploc_bytes(int cap) {
  return (size_t)cap * (4 + 4 + 32 + 32 + 4 + 4 + 4) + (size_t)cap * sizeof(BvhNode) + 4096;
}

PlocWork ploc_carve(void* base, int cap) {
  char* p = (char*)base;
  PlocWork w;
  w.ba = (float4*)p; p += (size_t)cap * 32;
  w.bb = (float4*)p; p += (size_t)cap * 32;
  w.nodes = (BvhNode*)p; p += (size_t)cap * sizeof(BvhNode);
  w.ca = (int*)p; p += (size_t)cap * 4;
  w.cb = (int*)p; p += (size_t)cap * 4;
  w.nn = (int*)p; p += (size_t)cap * 4;
  w.valid = (int*)p; p += (size_t)cap * 4;
  w.pos = (int*)p; p += (size_t)cap * 4;
  w.cnt = (int*)p; p += 64;
  w.alloc = (int*)p;
  return w;
}

__device__ __forceinline__ float union_area(float4 alo, float4 ahi, float4 blo, float4 bhi) {
  const float x = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
  const float y = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
  const float z = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
  return x * y + y * z + z * x;
}

__global__ void k_ploc_init(const float4* __restrict__ tris, int n, int* __restrict__ ca,
                            float4* __restrict__ ba, int* __restrict__ cnt,
                            int* __restrict__ alloc) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i == 0) {
    cnt[0] = n;
    cnt[1] = n;
    *alloc = 0;
  }
  if (i >= n) return;
  const float4 a = tris[3 * i], b = tris[3 * i + 1], c = tris[3 * i + 2];
  ca[i] = ~i;
  ba[2 * i] = make_float4(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)),
                          fminf(a.z, fminf(b.z, c.z)), 0.f);
  ba[2 * i + 1] = make_float4(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)),
                              fmaxf(a.z, fmaxf(b.z, c.z)), 0.f);
}

__global__ void k_ploc_nn(const float4* __restrict__ ba, const int* __restrict__ cnt,
                          int* __restrict__ nn) {
  const int count = *cnt;
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= count || count <= 1) return;
  const float4 lo = ba[2 * i], hi = ba[2 * i + 1];
  int best = -1;
  float best_a = INFINITY;
  const int j0 = i - kPlocRadius < 0 ? 0 : i - kPlocRadius;
  const int j1 = i + kPlocRadius > count - 1 ? count - 1 : i + kPlocRadius;
  for (int j = j0; j <= j1; ++j) {
    if (j == i) continue;
    const float a = union_area(lo, hi, ba[2 * j], ba[2 * j + 1]);
    if (a < best_a) {  // ties keep the smaller index (j ascends)
      best_a = a;
      best = j;
    }
  }
  nn[i] = best;
}

__global__ void k_ploc_merge(int* __restrict__ ca, float4* __restrict__ ba,
                             const int* __restrict__ nn, const int* __restrict__ cnt,
                             int* __restrict__ valid, int n_max, BvhNode* __restrict__ nodes,
                             int* __restrict__ alloc, int n) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n_max) return;
  const int count = *cnt;
  if (i >= count) {
    valid[i] = 0;
    return;
  }
  if (count <= 1) {
    valid[i] = 1;
    return;
  }
  const int j = nn[i];
  if (j < 0 || j >= count || nn[j] != i) {  // not a mutual pair: the cluster survives
    valid[i] = 1;
    return;
  }
  if (i > j) {  // the smaller index of the pair builds the node and keeps its slot
    valid[i] = 0;
    return;
  }
  const int node = (n - 2) - atomicAdd(alloc, 1);
  if (node < 0) {  // more merges than a binary tree has nodes: cannot happen; drop the pair
    valid[i] = 1;
    return;
  }
  const float4 l0 = ba[2 * i], l1 = ba[2 * i + 1], r0 = ba[2 * j], r1 = ba[2 * j + 1];
  BvhNode nd;
  nd.a = make_float4(l0.x, l0.y, l0.z, l1.x);
  nd.b = make_float4(l1.y, l1.z, r0.x, r0.y);
  nd.c = make_float4(r0.z, r1.x, r1.y, r1.z);
  nd.d = make_int4(ca[i], ca[j], 0, 0);
  nodes[node] = nd;
  ca[i] = node;
  ba[2 * i] = make_float4(fminf(l0.x, r0.x), fminf(l0.y, r0.y), fminf(l0.z, r0.z), 0.f);
  ba[2 * i + 1] = make_float4(fmaxf(l1.x, r1.x), fmaxf(l1.y, r1.y), fmaxf(l1.z, r1.z), 0.f);
  valid[i] = 1;
}

__global__ void k_ploc_scatter(const int* __restrict__ ca, const float4* __restrict__ ba,
                               const int* __restrict__ valid, const int* __restrict__ pos,
                               const int* __restrict__ cnt, int* __restrict__ cb,
                               float4* __restrict__ bb) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= *cnt || !valid[i]) return;
  const int k = pos[i];
  cb[k] = ca[i];
  bb[2 * k] = ba[2 * i];
  bb[2 * k + 1] = ba[2 * i + 1];
}

// The remaining rounds in one workgroup: the same nearest-neighbour, merge and in-order
// compaction steps, separated by workgroup barriers instead of launches.
constexpr int kFinBS = 1024;

__global__ void __launch_bounds__(kFinBS) k_ploc_finish(int* __restrict__ ca,
                                                        float4* __restrict__ ba,
                                                        int* __restrict__ cb,
                                                        float4* __restrict__ bb,
                                                        int* __restrict__ nn,
                                                        int* __restrict__ valid,
                                                        int* __restrict__ cnt,
                                                        BvhNode* __restrict__ nodes,
                                                        int* __restrict__ alloc, int n) {
  __shared__ int s_count;
  if (threadIdx.x == 0) s_count = *cnt;
  __syncthreads();
  int* A = ca;
  float4* AB = ba;
  int* B = cb;
  float4* BB = bb;
  for (int round = 0; round < kPlocFinishRounds; ++round) {
    const int count = s_count;
    if (count <= 1) break;
    for (int i = threadIdx.x; i < count; i += kFinBS) {
      const float4 lo = AB[2 * i], hi = AB[2 * i + 1];
      int best = -1;
      float best_a = INFINITY;
      const int j0 = i - kPlocRadius < 0 ? 0 : i - kPlocRadius;
      const int j1 = i + kPlocRadius > count - 1 ? count - 1 : i + kPlocRadius;
      for (int j = j0; j <= j1; ++j) {
        if (j == i) continue;
        const float a = union_area(lo, hi, AB[2 * j], AB[2 * j + 1]);
        if (a < best_a) {
          best_a = a;
          best = j;
        }
      }
      nn[i] = best;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < count; i += kFinBS) {
      const int j = nn[i];
      if (j < 0 || j >= count || nn[j] != i) {
        valid[i] = 1;
        continue;
      }
      if (i > j) {
        valid[i] = 0;
        continue;
      }
      const int node = (n - 2) - atomicAdd(alloc, 1);
      if (node < 0) {
        valid[i] = 1;
        continue;
      }
      const float4 l0 = AB[2 * i], l1 = AB[2 * i + 1], r0 = AB[2 * j], r1 = AB[2 * j + 1];
      BvhNode nd;
      nd.a = make_float4(l0.x, l0.y, l0.z, l1.x);
      nd.b = make_float4(l1.y, l1.z, r0.x, r0.y);
      nd.c = make_float4(r0.z, r1.x, r1.y, r1.z);
      nd.d = make_int4(A[i], A[j], 0, 0);
      nodes[node] = nd;
      A[i] = node;
      AB[2 * i] = make_float4(fminf(l0.x, r0.x), fminf(l0.y, r0.y), fminf(l0.z, r0.z), 0.f);
      AB[2 * i + 1] = make_float4(fmaxf(l1.x, r1.x), fmaxf(l1.y, r1.y), fmaxf(l1.z, r1.z), 0.f);
      valid[i] = 1;
    }
    __syncthreads();
    int carry = 0;
    for (int base = 0; base < count; base += kFinBS) {
      const int i = base + threadIdx.x;
      const int v = i < count ? valid[i] : 0;
      int tot;
      const int ex = petch::block_excl_scan<kFinBS>(v, tot);
      if (v) {
        const int k = carry + ex;
        B[k] = A[i];
        BB[2 * k] = AB[2 * i];
        BB[2 * k + 1] = AB[2 * i + 1];
      }
      carry += tot;
    }
    __syncthreads();
    if (threadIdx.x == 0) s_count = carry;
    __syncthreads();
    int* ti = A; A = B; B = ti;
    float4* tf = AB; AB = BB; BB = tf;
  }
  if (threadIdx.x == 0) *cnt = s_count;
}

// Adopts the clustered tree when it is complete (one cluster left after n-1 merges).
__global__ void k_ploc_commit(const BvhNode* __restrict__ src, const int* __restrict__ cnt,
                              const int* __restrict__ alloc, int n,
                              BvhNode* __restrict__ dst) {
  if (*cnt != 1 || *alloc != n - 1) return;
  for (int k = blockIdx.x * kBS + threadIdx.x; k < n - 1; k += gridDim.x * kBS) dst[k] = src[k];
}

__global__ void k_query_ray(BvhView b, const float* __restrict__ o, const float* __restrict__ d,
                            int n, float tmax, int* __restrict__ face, float* __restrict__ tout) {
  int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  petch::Hit h = petch::bvh_closest_hit(b, petch::v3(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                                        petch::v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]), tmax);
  face[i] = h.face;
  if (tout) tout[i] = h.face >= 0 ? h.t : -1.f;
}

__global__ void k_query_point(BvhView b, const float* __restrict__ p, int n, float maxd,
                              int* __restrict__ face, float* __restrict__ dist) {
  int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  petch::PointHit h =
      petch::bvh_closest_point(b, petch::v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), maxd);
  face[i] = h.face;
  if (dist) dist[i] = h.face >= 0 ? sqrtf(h.d2) : -1.f;
}

}  // namespace

BvhView petch_bvh_view(const petch_bvh* b) {
  BvhView v;
  v.nodes = (const BvhNode*)b->nodes;
  v.wide = (const BvhWide*)b->wide;
  v.tris = (const float4*)b->tris;
  v.nf = b->nf;
  v.err = b->work ? carve(b->work, b->cap_f).err : nullptr;
  v.root = (b->work && b->nf > 1) ? carve(b->work, b->cap_f).root : nullptr;
  return v;
}

extern "C" int petch_bvh_create(petch_ctx* ctx, petch_bvh** out) {
  PETCH_ARG(ctx, ctx && out, "bvh_create: null pointer");
  petch_bvh* b = (petch_bvh*)calloc(1, sizeof(petch_bvh));
  if (!b) return petch_set_error(ctx, PETCH_E_HIP, "bvh_create: out of host memory");
  b->device = ctx->device;
  b->quality = -1;
  *out = b;
  return PETCH_OK;
}

extern "C" int petch_bvh_set_quality(petch_ctx* ctx, petch_bvh* b, int quality) {
  PETCH_ARG(ctx, ctx && b, "bvh_set_quality: null pointer");
  PETCH_ARG(ctx, quality >= -1 && quality <= 2, "bvh_set_quality: quality %d not in -1..2",
            quality);
  b->quality = quality;
  return PETCH_OK;
}

extern "C" int petch_bvh_destroy(petch_ctx* ctx, petch_bvh* b) {
  if (!b) return PETCH_OK;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  if (b->nodes) (void)hipFree(b->nodes);
  if (b->wide) (void)hipFree(b->wide);
  if (b->tris) (void)hipFree(b->tris);
  if (b->work) (void)hipFree(b->work);
  if (b->aux) (void)hipFree(b->aux);
  free(b);
  return PETCH_OK;
}

namespace {

// PLOC rounds over the leaves of b (tris already in Morton order); the result replaces b's
// Morton tree only when it is complete.
int ploc_build(petch_ctx* ctx, petch_bvh* b, int nf) {
  const size_t need = ploc_bytes(nf);
  if (b->aux_cap < need) {
    PETCH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (b->aux) (void)hipFree(b->aux);
    b->aux = nullptr;
    b->aux_cap = 0;
    const size_t want = need + need / 2;
    PETCH_HIP(ctx, hipMalloc(&b->aux, want));
    b->aux_cap = want;
  }
  PlocWork p = ploc_carve(b->aux, nf);
  const unsigned g = petch_blocks(nf, kBS);
  hipLaunchKernelGGL(k_ploc_init, dim3(g), dim3(kBS), 0, ctx->stream, (const float4*)b->tris, nf,
                     p.ca, p.ba, p.cnt, p.alloc);
  PETCH_LAUNCH_CHECK(ctx);
  int *ca = p.ca, *cb = p.cb;
  float4 *ba = p.ba, *bb = p.bb;
  for (int r = 0; r < kPlocRounds; ++r) {
    int* cur = p.cnt + (r & 1);
    int* nxt = p.cnt + ((r + 1) & 1);
    hipLaunchKernelGGL(k_ploc_nn, dim3(g), dim3(kBS), 0, ctx->stream, ba, cur, p.nn);
    PETCH_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_ploc_merge, dim3(g), dim3(kBS), 0, ctx->stream, ca, ba, p.nn, cur,
                       p.valid, nf, p.nodes, p.alloc, nf);
    PETCH_LAUNCH_CHECK(ctx);
    int rc = petch_exclusive_scan(ctx, p.valid, p.pos, nf, nxt);
    if (rc != PETCH_OK) return rc;
    hipLaunchKernelGGL(k_ploc_scatter, dim3(g), dim3(kBS), 0, ctx->stream, ca, ba, p.valid,
                       p.pos, cur, cb, bb);
    PETCH_LAUNCH_CHECK(ctx);
    int* ti = ca; ca = cb; cb = ti;
    float4* tf = ba; ba = bb; bb = tf;
  }
  int* cnt = p.cnt + (kPlocRounds & 1);
  hipLaunchKernelGGL(k_ploc_finish, dim3(1), dim3(kFinBS), 0, ctx->stream, ca, ba, cb, bb, p.nn,
                     p.valid, cnt, p.nodes, p.alloc, nf);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_ploc_commit, dim3(g < 256 ? g : 256), dim3(kBS), 0, ctx->stream, p.nodes,
                     cnt, p.alloc, nf, (BvhNode*)b->nodes);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

}  // namespace

namespace {

// Grows the BVH's device buffers to hold nf faces (they are never shrunk); a BVH that was never
// built gets a minimal set, so its error flag is valid. A regrowth waits for the stream and
// frees and allocates four buffers (hundreds of microseconds with the device idle), and an
// etching surface keeps growing, so the new capacity is twice the request.
int ensure_capacity(petch_ctx* ctx, petch_bvh* b, int nf) {
  if (nf > b->cap_f) {
    PETCH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (b->nodes) (void)hipFree(b->nodes);
    if (b->wide) (void)hipFree(b->wide);
    if (b->tris) (void)hipFree(b->tris);
    if (b->work) (void)hipFree(b->work);
    b->nodes = b->wide = b->tris = b->work = nullptr;
    int cap = nf > (0x7fffffff - 1024) / 2 ? nf : 2 * nf + 1024;
    PETCH_HIP(ctx, hipMalloc(&b->nodes, (size_t)cap * sizeof(BvhNode)));
    PETCH_HIP(ctx, hipMalloc(&b->wide, (size_t)cap * sizeof(BvhWide)));
    PETCH_HIP(ctx, hipMalloc(&b->tris, (size_t)cap * 48));
    PETCH_HIP(ctx, hipMalloc(&b->work, work_bytes(cap)));
    b->cap_f = cap;
  }
  if (b->cap_f == 0) {  // never built: keep a valid error flag
    int cap = 256;
    PETCH_HIP(ctx, hipMalloc(&b->nodes, (size_t)cap * sizeof(BvhNode)));
    PETCH_HIP(ctx, hipMalloc(&b->wide, (size_t)cap * sizeof(BvhWide)));
    PETCH_HIP(ctx, hipMalloc(&b->tris, (size_t)cap * 48));
    PETCH_HIP(ctx, hipMalloc(&b->work, work_bytes(cap)));
    b->cap_f = cap;
  }
  return PETCH_OK;
}

// Everything of a build up to the refit: buffers, centroids, Morton sort, leaves, Karras nodes
// and the cleared refit flags. Returns the refit's item (n = nf; nothing to refit when nf < 2).
int build_prepare(petch_ctx* ctx, petch_bvh* b, const float* verts, const int* faces, int nv,
                  int nf, RefitItem* refit) {
  PETCH_ARG(ctx, ctx && b, "bvh_build: null pointer");
  PETCH_ARG(ctx, nf >= 0 && nv >= 0, "bvh_build: negative size");
  PETCH_ARG(ctx, nf == 0 || (verts && faces), "bvh_build: null mesh");
  b->nv = nv;
  b->nf = 0;
  refit->n = 0;
  const int rc0 = ensure_capacity(ctx, b, nf);
  if (rc0 != PETCH_OK) return rc0;
  Work w = carve(b->work, b->cap_f);
  PETCH_HIP(ctx, hipMemsetAsync(w.err, 0, sizeof(int), ctx->stream));
  if (nf == 0) return PETCH_OK;
  unsigned g = petch_blocks(nf, kBS);
  hipLaunchKernelGGL(k_bvh_prep, dim3(g), dim3(kBS), 0, ctx->stream, verts, faces, nf, w.cen,
                     w.part);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_bvh_bounds, dim3(1), dim3(kBS), 0, ctx->stream, w.part, (int)g, w.scene);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_bvh_morton, dim3(g), dim3(kBS), 0, ctx->stream, w.cen, nf, w.scene, w.keys,
                     w.vals);
  PETCH_LAUNCH_CHECK(ctx);
  int rc = petch_radix_sort_pairs(ctx, w.keys, w.vals, w.ktmp, w.vtmp, nf, 30);
  if (rc != PETCH_OK) return rc;
  hipLaunchKernelGGL(k_bvh_leaves, dim3(g), dim3(kBS), 0, ctx->stream, verts, faces, w.vals, nf,
                     (float4*)b->tris);
  PETCH_LAUNCH_CHECK(ctx);
  if (nf > 1) {
    hipLaunchKernelGGL(k_bvh_karras, dim3(petch_blocks(nf - 1, kBS)), dim3(kBS), 0, ctx->stream,
                       w.keys, nf, (BvhNode*)b->nodes, w.pint, w.pleaf);
    PETCH_LAUNCH_CHECK(ctx);
    PETCH_HIP(ctx, hipMemsetAsync(w.flags, 0, (size_t)nf * sizeof(int), ctx->stream));
    *refit = RefitItem{(const float4*)b->tris, (BvhNode*)b->nodes, w.pint, w.pleaf, w.flags,
                       w.root, nf};
  }
  return PETCH_OK;
}

// The rest of a build after the refit: the clustered tree where the quality asks for it, then
// the 4-wide nodes the ray queries walk.
int build_finish(petch_ctx* ctx, petch_bvh* b, int nf) {
  if (nf > 1) {
    Work w = carve(b->work, b->cap_f);
    const int q = b->quality >= 0 ? b->quality : (nf >= kWideSahMinFaces ? 2 : 0);
    if (q >= 2) {
      int rc2 = ploc_build(ctx, b, nf);
      if (rc2 != PETCH_OK) return rc2;
    }
    if (q >= 1) {
      hipLaunchKernelGGL(k_bvh_wide_sah, dim3(1), dim3(kWideBS), 0, ctx->stream,
                         (const BvhNode*)b->nodes, nf, (BvhWide*)b->wide, (int*)w.ktmp,
                         (int*)w.vtmp, w.err);
    } else {
      hipLaunchKernelGGL(k_bvh_wide, dim3(petch_blocks(nf - 1, kBS)), dim3(kBS), 0, ctx->stream,
                         (const BvhNode*)b->nodes, w.pint, nf, (BvhWide*)b->wide);
    }
    PETCH_LAUNCH_CHECK(ctx);
  }
  b->nf = nf;
  return PETCH_OK;
}

}  // namespace

extern "C" int petch_bvh_build(petch_ctx* ctx, petch_bvh* b, const float* verts, const int* faces,
                               int nv, int nf) {
  RefitItem r;
  int rc = build_prepare(ctx, b, verts, faces, nv, nf, &r);
  if (rc != PETCH_OK) return rc;
  if (r.n > 1) {
    hipLaunchKernelGGL(k_bvh_refit, dim3(petch_blocks(r.n, kBS)), dim3(kBS), 0, ctx->stream,
                       r.tris, r.nodes, r.pint, r.pleaf, r.flags, r.n, r.root);
    PETCH_LAUNCH_CHECK(ctx);
  }
  return build_finish(ctx, b, nf);
}

extern "C" int petch_bvh_build_batch(petch_ctx* ctx, int n, petch_bvh* const* bvhs,
                                     const float* const* verts, const int* const* faces,
                                     const int* nv, const int* nf) {
  PETCH_ARG(ctx, ctx && (n == 0 || (bvhs && verts && faces && nv && nf)),
            "bvh_build_batch: null pointer");
  PETCH_ARG(ctx, n >= 0, "bvh_build_batch: n < 0");
  if (n == 0) return PETCH_OK;
  for (int t = 0; t < n; ++t) {
    PETCH_ARG(ctx, bvhs[t], "bvh_build_batch: item %d has a null BVH", t);
    PETCH_ARG(ctx, nf[t] >= 0 && nv[t] >= 0, "bvh_build_batch: item %d has a negative size", t);
    PETCH_ARG(ctx, nf[t] == 0 || (verts[t] && faces[t]), "bvh_build_batch: item %d has a null "
              "mesh", t);
  }
  // Morton trees of two or more faces take the batched stages; the others (clustered or
  // area-collapsed trees, meshes under two faces) are prepared one by one
  std::vector<RefitItem> refit;
  std::vector<BuildItem> fast;
  std::vector<petch_sort_item> sorts;
  std::vector<int> fast_t;
  long long fb = 0, nb = 0, tiles = 0, tab = 0;
  for (int t = 0; t < n; ++t) {
    petch_bvh* b = bvhs[t];
    const int q = b->quality >= 0 ? b->quality : (nf[t] >= kWideSahMinFaces ? 2 : 0);
    if (q != 0 || nf[t] < 2) {
      RefitItem r;
      const int rc = build_prepare(ctx, b, verts[t], faces[t], nv[t], nf[t], &r);
      if (rc != PETCH_OK) return rc;
      if (r.n > 1) refit.push_back(r);
      continue;
    }
    b->nv = nv[t];
    b->nf = 0;
    const int rc = ensure_capacity(ctx, b, nf[t]);
    if (rc != PETCH_OK) return rc;
    BuildItem it;
    it.verts = verts[t];
    it.faces = faces[t];
    it.nf = nf[t];
    it.nfb = (int)petch_blocks(nf[t], kBS);
    it.w = carve(b->work, b->cap_f);
    it.nodes = (BvhNode*)b->nodes;
    it.wide = (BvhWide*)b->wide;
    it.tris = (float4*)b->tris;
    it.fblk0 = fb;
    it.nblk0 = nb;
    fb += it.nfb;
    nb += petch_blocks(nf[t] - 1, kBS);
    petch_sort_item si;
    si.keys = it.w.keys;
    si.vals = it.w.vals;
    si.ktmp = it.w.ktmp;
    si.vtmp = it.w.vtmp;
    si.n = nf[t];
    si.nblocks = petch_sort_tiles(nf[t]);
    si.tblk0 = tiles;
    si.hist = si.offs = nullptr;
    tiles += si.nblocks;
    tab += 2LL * 256 * si.nblocks;
    fast.push_back(it);
    sorts.push_back(si);
    fast_t.push_back(t);
    refit.push_back(RefitItem{it.tris, it.nodes, it.w.pint, it.w.pleaf, it.w.flags, it.w.root,
                              it.nf});
  }
  PETCH_ARG(ctx, fb <= 0x7fffffffLL && tiles <= 0x7fffffffLL,
            "bvh_build_batch: too many faces for one launch");
  const int m = (int)fast.size();
  if (m > 0) {
    int* tabs = (int*)petch_scratch(ctx, SCR_SORT_MANY, (size_t)tab * sizeof(int));
    if (!tabs) return PETCH_E_HIP;
    long long off = 0;
    for (auto& si : sorts) {
      si.hist = tabs + off;
      si.offs = tabs + off + 256LL * si.nblocks;
      off += 2LL * 256 * si.nblocks;
    }
  }
  // one table upload for every stage: build items, sort items, refit items
  const size_t sz_b = fast.size() * sizeof(BuildItem), sz_s = sorts.size() * sizeof(petch_sort_item);
  const size_t off_s = (sz_b + 255) & ~(size_t)255, off_r = off_s + ((sz_s + 255) & ~(size_t)255);
  const size_t total = off_r + refit.size() * sizeof(RefitItem);
  if (refit.empty()) {
    for (int t = 0; t < n; ++t) {
      const int rc = build_finish(ctx, bvhs[t], nf[t]);
      if (rc != PETCH_OK) return rc;
    }
    return PETCH_OK;
  }
  std::vector<char> blob(total);
  if (sz_b) memcpy(blob.data(), fast.data(), sz_b);
  if (sz_s) memcpy(blob.data() + off_s, sorts.data(), sz_s);
  memcpy(blob.data() + off_r, refit.data(), refit.size() * sizeof(RefitItem));
  char* d = (char*)petch_upload_table(ctx, blob.data(), total);
  if (!d) return PETCH_E_HIP;
  const BuildItem* db = (const BuildItem*)d;
  const petch_sort_item* ds = (const petch_sort_item*)(d + off_s);
  const RefitItem* dr = (const RefitItem*)(d + off_r);
  if (m > 0) {
    hipLaunchKernelGGL(k_bvh_prep_many, dim3((unsigned)fb), dim3(kBS), 0, ctx->stream, db, m);
    PETCH_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_bvh_bounds_many, dim3((unsigned)m), dim3(kBS), 0, ctx->stream, db);
    PETCH_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_bvh_morton_many, dim3((unsigned)fb), dim3(kBS), 0, ctx->stream, db, m);
    PETCH_LAUNCH_CHECK(ctx);
    const int rc = petch_radix_sort_pairs_many(ctx, ds, m, tiles);
    if (rc != PETCH_OK) return rc;
    hipLaunchKernelGGL(k_bvh_leaves_many, dim3((unsigned)fb), dim3(kBS), 0, ctx->stream, db, m);
    PETCH_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_bvh_karras_many, dim3((unsigned)nb), dim3(kBS), 0, ctx->stream, db, m);
    PETCH_LAUNCH_CHECK(ctx);
  }
  hipLaunchKernelGGL(k_bvh_refit_wg, dim3((unsigned)refit.size()), dim3(kRefitWG), 0, ctx->stream,
                     dr);
  PETCH_LAUNCH_CHECK(ctx);
  if (m > 0) {
    hipLaunchKernelGGL(k_bvh_wide_many, dim3((unsigned)nb), dim3(kBS), 0, ctx->stream, db, m);
    PETCH_LAUNCH_CHECK(ctx);
    for (int k = 0; k < m; ++k) bvhs[fast_t[k]]->nf = nf[fast_t[k]];
  }
  // the others: their clustered tree / area-driven collapse, one by one
  std::vector<char> done((size_t)n, 0);
  for (int k = 0; k < m; ++k) done[(size_t)fast_t[k]] = 1;
  for (int t = 0; t < n; ++t) {
    if (done[(size_t)t]) continue;
    const int rc = build_finish(ctx, bvhs[t], nf[t]);
    if (rc != PETCH_OK) return rc;
  }
  return PETCH_OK;
}

extern "C" int petch_bvh_query_ray(petch_ctx* ctx, const petch_bvh* b, const float* orig,
                                   const float* dir, int n, float tmax, int* face, float* t) {
  PETCH_ARG(ctx, ctx && b && orig && dir && face, "bvh_query_ray: null pointer");
  PETCH_ARG(ctx, n >= 0, "bvh_query_ray: n < 0");
  if (n == 0) return PETCH_OK;
  hipLaunchKernelGGL(k_query_ray, dim3(petch_blocks(n, kBS)), dim3(kBS), 0, ctx->stream,
                     petch_bvh_view(b), orig, dir, n, tmax, face, t);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_bvh_query_point(petch_ctx* ctx, const petch_bvh* b, const float* pts, int n,
                                     float max_dist, int* face, float* dist) {
  PETCH_ARG(ctx, ctx && b && pts && face, "bvh_query_point: null pointer");
  PETCH_ARG(ctx, n >= 0, "bvh_query_point: n < 0");
  if (n == 0) return PETCH_OK;
  hipLaunchKernelGGL(k_query_point, dim3(petch_blocks(n, kBS)), dim3(kBS), 0, ctx->stream,
                     petch_bvh_view(b), pts, n, max_dist, face, dist);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_bvh_status(petch_ctx* ctx, const petch_bvh* b, int* err_host) {
  PETCH_ARG(ctx, ctx && b && err_host, "bvh_status: null pointer");
  *err_host = 0;
  if (!b->work) return PETCH_OK;
  int* h = (int*)petch_pinned(ctx, sizeof(int));
  if (!h) return PETCH_E_HIP;
  PETCH_HIP(ctx, hipMemcpyAsync(h, carve(b->work, b->cap_f).err, sizeof(int),
                                hipMemcpyDeviceToHost, ctx->stream));
  PETCH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *err_host = *h;
  return PETCH_OK;
}

extern "C" int petch_bvh_status_async(petch_ctx* ctx, const petch_bvh* b, int* err_dev) {
  PETCH_ARG(ctx, ctx && b && err_dev, "bvh_status_async: null pointer");
  if (!b->work) {
    PETCH_HIP(ctx, hipMemsetAsync(err_dev, 0, sizeof(int), ctx->stream));
    return PETCH_OK;
  }
  PETCH_HIP(ctx, hipMemcpyAsync(err_dev, carve(b->work, b->cap_f).err, sizeof(int),
                                hipMemcpyDeviceToDevice, ctx->stream));
  return PETCH_OK;
}

extern "C" int petch_bvh_root_bounds(petch_ctx* ctx, const petch_bvh* b, float* lohi_host) {
  PETCH_ARG(ctx, ctx && b && lohi_host, "bvh_root_bounds: null pointer");
  PETCH_ARG(ctx, b->work && b->nf > 1, "bvh_root_bounds: BVH has no internal nodes");
  PETCH_HIP(ctx, hipMemcpyAsync(lohi_host, carve(b->work, b->cap_f).root, 6 * sizeof(float),
                                hipMemcpyDeviceToHost, ctx->stream));
  PETCH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PETCH_OK;
}
