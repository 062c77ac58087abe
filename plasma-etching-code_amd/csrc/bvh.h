// Device-side view of the triangle LBVH and its two queries:
//   closest two-sided ray hit (wp.mesh_query_ray semantics, threed.py:306,366,524,569,629,916)
//   closest face to a point   (wp.mesh_query_point_no_sign, threed.py:1328,1626)
//
// Layout in HBM:
//   nodes: (nf-1) internal nodes of 64 B = both children's AABBs + child references
//          (>= 0 internal node, < 0 leaf ~slot); node 0 is the root.
//   tris:  nf leaves of 48 B in Morton order: v0 (w = face id bits), v1, v2 as float4, so a
//          leaf is three 16-byte loads and carries the face id the flux deposits into.
//   wide:  4-wide nodes of 128 B collapsed from the binary tree, each stored at the index of
//          the binary node it replaces (children: up to four descendants; a leaf stays a leaf).
//          Ray queries walk the wide nodes, which halves the chain of dependent node fetches;
//          point queries walk the binary nodes.
// Traversal keeps a short per-lane stack of far children with their entry distances (the
// nearest child is descended first, and a popped entry beyond the best hit so far is dropped
// unfetched); a stack overflow drops the far subtree and raises error flag 1 instead of writing
// out of bounds; an out-of-range reference or a step count beyond a tree's ends the query with
// flag 2.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace petch {

constexpr int kBvhStack = 64;
constexpr float kBoxSlack = 1.00000024f;  // 1 + 2 ulp

struct __align__(16) BvhNode {
  float4 a;  // lmin.x lmin.y lmin.z lmax.x
  float4 b;  // lmax.y lmax.z rmin.x rmin.y
  float4 c;  // rmin.z rmax.x rmax.y rmax.z
  int4 d;    // left, right, -, -
};

// Four children's boxes, one child per float4 lane, and their references: >= 0 a wide node
// (indexed like its binary node), < 0 a leaf ~slot, kWideEmpty an unused lane. An unused lane's
// box has every bound at +inf: each slab then gives +inf (or -inf) at both planes, so the slab
// test misses it for every ray without looking at the reference.
struct __align__(16) BvhWide {
  float4 lox, hix, loy, hiy, loz, hiz;
  int4 ref;
  int4 pad;
};
constexpr int kWideEmpty = (int)0x80000000;

struct BvhView {
  const BvhNode* nodes;
  const BvhWide* wide;
  const float4* tris;
  int nf;
  int* err;           // device flag: set to 1 on traversal stack overflow (may be null)
  const float* root;  // bounds of all triangles (lo xyz, hi xyz); null below two faces
};

struct Vec3 {
  float x, y, z;
};

__device__ __forceinline__ Vec3 v3(float x, float y, float z) { return Vec3{x, y, z}; }
__device__ __forceinline__ Vec3 operator+(Vec3 a, Vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ Vec3 operator-(Vec3 a, Vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ Vec3 operator*(float s, Vec3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ Vec3 operator-(Vec3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ Vec3 cross(Vec3 a, Vec3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length(Vec3 a) { return sqrtf(dot(a, a)); }
// x/|x|, or the zero vector for a zero-length x
__device__ __forceinline__ Vec3 normalize(Vec3 a) {
  float l = length(a);
  return l > 0.f ? (1.f / l) * a : v3(0.f, 0.f, 0.f);
}
__device__ __forceinline__ float comp(Vec3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// Per-ray constants of the watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013):
// kz is the direction's largest component and (kx, ky) = (kz+1, kz+2) mod 3. The paper swaps
// kx and ky for a negative d[kz] to keep the winding; the test here is two-sided, and the swap
// only negates U, V, W, det and T exactly, so it is left out.
struct RayPre {
  Vec3 o, inv;
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

// Moves the ray's origin (a periodic wrap keeps the direction and its constants).
__device__ __forceinline__ void ray_move(RayPre& r, Vec3 o) { r.o = o; }

__device__ __forceinline__ RayPre ray_pre(Vec3 o, Vec3 d) {
  RayPre r;
  r.o = o;
  // reciprocal of a zero component is taken of +-1e-20 so slab products stay finite (no 0*inf)
  r.inv = v3(1.f / (fabsf(d.x) > 1e-20f ? d.x : copysignf(1e-20f, d.x)),
             1.f / (fabsf(d.y) > 1e-20f ? d.y : copysignf(1e-20f, d.y)),
             1.f / (fabsf(d.z) > 1e-20f ? d.z : copysignf(1e-20f, d.z)));
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  r.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  r.kx = r.kz == 2 ? 0 : r.kz + 1;
  r.ky = r.kx == 2 ? 0 : r.kx + 1;
  const float dz = comp(d, r.kz);
  r.Sx = comp(d, r.kx) / dz;
  r.Sy = comp(d, r.ky) / dz;
  r.Sz = 1.f / dz;
  return r;
}

// Two-sided watertight test; on a hit with 0 < t < tbest returns true and sets t.
__device__ __forceinline__ bool ray_tri(const RayPre& r, Vec3 p0, Vec3 p1, Vec3 p2, float tbest,
                                        float& t) {
#pragma clang fp contract(off)
  Vec3 A = p0 - r.o, B = p1 - r.o, C = p2 - r.o;
  float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
  float Ax = comp(A, r.kx) - r.Sx * Akz, Ay = comp(A, r.ky) - r.Sy * Akz;
  float Bx = comp(B, r.kx) - r.Sx * Bkz, By = comp(B, r.ky) - r.Sy * Bkz;
  float Cx = comp(C, r.kx) - r.Sx * Ckz, Cy = comp(C, r.ky) - r.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    // edge hit: re-evaluate the edge functions in double so shared edges stay closed
    U = (float)((double)Cx * By - (double)Cy * Bx);
    V = (float)((double)Ax * Cy - (double)Ay * Cx);
    W = (float)((double)Bx * Ay - (double)By * Ax);
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
  float det = U + V + W;
  if (det == 0.f) return false;
  float T = U * (r.Sz * Akz) + V * (r.Sz * Bkz) + W * (r.Sz * Ckz);
  if (det < 0.f) {
    T = -T;
    det = -det;
  }
  if (!(T > 0.f) || !(T < tbest * det)) return false;
  t = T / det;
  return t > 0.f && t < tbest;
}

// Slab constants of a ray for the fused box test. A box plane p on axis a is crossed at
// t = (p - o_a) * inv_a; the test evaluates it as one fma, fma(p, inv_a, -o_a * inv_a), on the
// near plane (lo for a positive direction, hi for a negative one) and the far plane of each
// axis. The product o_a * inv_a is rounded once, which can shift t by up to 2^-24 |o_a inv_a|,
// so the near offset is taken 2^-21 |o_a inv_a| lower and the far offset as much higher: the
// fused test keeps every box that (p - o_a) * inv_a evaluated exactly keeps, and the fma's
// own rounding is covered by the relative slack kBoxSlack. near_x etc. are the byte offsets of
// the near planes in a BvhWide (lox at 0, hix at 16, loy 32, hiy 48, loz 64, hiz 80).
struct Slabs {
  float nx, ny, nz, fx, fy, fz;
  int near_x, near_y, near_z;
};

__device__ __forceinline__ Slabs ray_slabs(const RayPre& r) {
  const float ox = r.o.x * r.inv.x, oy = r.o.y * r.inv.y, oz = r.o.z * r.inv.z;
  const float ex = fabsf(ox) * 4.76837158e-7f, ey = fabsf(oy) * 4.76837158e-7f,
              ez = fabsf(oz) * 4.76837158e-7f;
  Slabs s;
  s.nx = -ox - ex;
  s.ny = -oy - ey;
  s.nz = -oz - ez;
  s.fx = -ox + ex;
  s.fy = -oy + ey;
  s.fz = -oz + ez;
  s.near_x = r.inv.x < 0.f ? 16 : 0;
  s.near_y = r.inv.y < 0.f ? 48 : 32;
  s.near_z = r.inv.z < 0.f ? 80 : 64;
  return s;
}

// Entry distance of the ray into the box with near planes (nx, ny, nz) and far planes
// (fx, fy, fz), clipped to [0, tbest]; kMiss when the ray misses it.
constexpr float kMiss = 3.0e38f;   // entry distance of a box the ray misses
constexpr float kHitLim = 1.0e38f;
__device__ __forceinline__ float slab_entry(const RayPre& r, const Slabs& s, float nx, float ny,
                                            float nz, float fx, float fy, float fz, float tbest) {
  const float tn = fmaxf(fmaxf(__builtin_fmaf(nx, r.inv.x, s.nx), __builtin_fmaf(ny, r.inv.y, s.ny)),
                         fmaxf(__builtin_fmaf(nz, r.inv.z, s.nz), 0.f));
  const float tf = fminf(fminf(__builtin_fmaf(fx, r.inv.x, s.fx), __builtin_fmaf(fy, r.inv.y, s.fy)),
                         fminf(__builtin_fmaf(fz, r.inv.z, s.fz), tbest));
  // tf widened by 2 ulp so rounding never culls a box the ray touches (conservative)
  return tn <= tf * kBoxSlack ? tn : kMiss;
}

// Whether a ray with 0 < t < 1e6 can meet the box of all the triangles at all: the fused slab
// test above on the root bounds, as conservative as the tests of the root's children (their
// union is the same box). A ray it rejects misses every triangle, so a query for it would visit
// the root node and end in a miss.
__device__ __forceinline__ bool root_may_hit(const BvhView& b, const RayPre& r) {
  if (!b.root) return true;
  const float* q = b.root;
  const Slabs s = ray_slabs(r);
  const bool px = r.inv.x >= 0.f, py = r.inv.y >= 0.f, pz = r.inv.z >= 0.f;
  return slab_entry(r, s, px ? q[0] : q[3], py ? q[1] : q[4], pz ? q[2] : q[5],
                    px ? q[3] : q[0], py ? q[4] : q[1], pz ? q[5] : q[2], 1.0e6f) < kHitLim;
}

struct Hit {
  int face;   // original face id, -1 on a miss
  int slot;   // leaf slot
  float t;
};

__device__ __forceinline__ void leaf_test(const BvhView& b, const RayPre& r, int slot, Hit& h) {
  if (slot < 0 || slot >= b.nf) return;
  float4 q0 = b.tris[3 * slot], q1 = b.tris[3 * slot + 1], q2 = b.tris[3 * slot + 2];
  float t;
  if (ray_tri(r, v3(q0.x, q0.y, q0.z), v3(q1.x, q1.y, q1.z), v3(q2.x, q2.y, q2.z), h.t, t)) {
    h.t = t;
    h.face = __float_as_int(q0.w);
    h.slot = slot;
  }
}

__device__ __forceinline__ void sort2(float& ta, int& ca, float& tb, int& cb) {
  const bool sw = tb < ta;
  const float t = sw ? tb : ta;
  tb = sw ? ta : tb;
  ta = t;
  const int c = sw ? cb : ca;
  cb = sw ? ca : cb;
  ca = c;
}


// Per-lane traversal stack of (node, entry distance). The first L entries live in LDS, in this
// lane's column of a workgroup array with row stride S (lane-major, so the lanes of a wavefront
// hit distinct banks whatever their depths); deeper entries spill to a scratch array. With
// L = 0 the whole stack is in scratch.
template <int S, int L>
struct TravStack {
  int* lds_n;
  float* lds_t;
  int spill_n[kBvhStack - L];
  float spill_t[kBvhStack - L];
  __device__ __forceinline__ void put(int i, int n, float t) {
    if (L > 0 && i < L) {
      lds_n[i * S] = n;
      lds_t[i * S] = t;
    } else {
      spill_n[i - L] = n;
      spill_t[i - L] = t;
    }
  }
  __device__ __forceinline__ int node(int i) const {
    if (L > 0 && i < L) return lds_n[i * S];
    return spill_n[i - L];
  }
  static constexpr int kLds = L;
  // entry i of the LDS part (i < L)
  __device__ __forceinline__ float lds_dist(int i) const { return lds_t[i * S]; }
  __device__ __forceinline__ void lds_put(int i, int n, float t) {
    lds_n[i * S] = n;
    lds_t[i * S] = t;
  }
  __device__ __forceinline__ int lds_node(int i) const { return lds_n[i * S]; }
  __device__ __forceinline__ float dist(int i) const {
    if (L > 0 && i < L) return lds_t[i * S];
    return spill_t[i - L];
  }
};
using ScratchStack = TravStack<1, 0>;

// Closest-hit query as a loop of steps: trav_begin sets it up, each trav_step visits one wide
// node (its leaf children are tested on the spot) and returns true once the query is finished,
// with the closest hit in T.h. Trav holds only scalars so that it lives in registers; the ray
// constants stay with the caller (one copy, which a periodic wrap updates in place) and the
// stack is separate.
constexpr int kTravDone = -1;

struct Trav {
  Hit h;
  int node, sp, steps;
};

__device__ __forceinline__ void trav_begin(const BvhView& b, Trav& T, const RayPre& r,
                                           float tmax) {
  T.h = Hit{-1, -1, tmax};
  T.sp = 0;
  T.steps = 0;
  T.node = 0;
  if (b.nf == 1) leaf_test(b, r, 0, T.h);
  if (b.nf <= 1) T.node = kTravDone;
}

template <class Stack>
__device__ __forceinline__ void push_far(const BvhView& b, Stack& stk, int& sp, int node, float t) {
  if (sp < kBvhStack) {
    stk.put(sp, node, t);
    ++sp;
  } else if (b.err) {
    *b.err = 1;
  }
}

template <class Stack>
__device__ __forceinline__ bool trav_step(const BvhView& b, const RayPre& r, Trav& T, Stack& stk) {
  const int node = T.node;
  if (node == kTravDone) return true;
  // a tree visits each wide node at most once; the bound only guards a corrupt tree
  if (++T.steps > 2 * b.nf + 64 || node < 0 || node >= b.nf - 1) {
    if (b.err) *b.err = 2;
    T.node = kTravDone;
    return true;
  }
  Hit& h = T.h;
  // the near and far plane of each axis, picked by the direction's sign, for all four children
  const BvhWide* N = b.wide + node;
  const char* base = (const char*)N;
  const Slabs sl = ray_slabs(r);
  const float4 nx = *(const float4*)(base + sl.near_x), fx = *(const float4*)(base + (16 - sl.near_x));
  const float4 ny = *(const float4*)(base + sl.near_y), fy = *(const float4*)(base + (80 - sl.near_y));
  const float4 nz = *(const float4*)(base + sl.near_z), fz = *(const float4*)(base + (144 - sl.near_z));
  const int4 ref = N->ref;
  int c0 = ref.x, c1 = ref.y, c2 = ref.z, c3 = ref.w;
  // unused lanes carry a box at +inf that every ray misses (near +inf, or far -inf for a
  // negative direction), so all four tests run branch-free and the node's loads issue together
  float t0 = slab_entry(r, sl, nx.x, ny.x, nz.x, fx.x, fy.x, fz.x, h.t);
  float t1 = slab_entry(r, sl, nx.y, ny.y, nz.y, fx.y, fy.y, fz.y, h.t);
  float t2 = slab_entry(r, sl, nx.z, ny.z, nz.z, fx.z, fy.z, fz.z, h.t);
  float t3 = slab_entry(r, sl, nx.w, ny.w, nz.w, fx.w, fy.w, fz.w, h.t);
  // leaves are tested on the spot, in child order, as one loop over the lane's hit leaves (the
  // wavefront runs as many triangle tests as its busiest lane needs, not one per child slot)
  unsigned lm = (t0 < kHitLim && c0 < 0 ? 1u : 0u) | (t1 < kHitLim && c1 < 0 ? 2u : 0u) |
                (t2 < kHitLim && c2 < 0 ? 4u : 0u) | (t3 < kHitLim && c3 < 0 ? 8u : 0u);
  if (lm & 1u) t0 = kMiss;
  if (lm & 2u) t1 = kMiss;
  if (lm & 4u) t2 = kMiss;
  if (lm & 8u) t3 = kMiss;
  while (lm) {
    const int c = (lm & 1u) ? c0 : (lm & 2u) ? c1 : (lm & 4u) ? c2 : c3;
    lm &= lm - 1u;
    leaf_test(b, r, ~c, h);
  }
  // inner children that now start beyond the best hit are dropped
  const float lim = h.t * kBoxSlack;
  if (!(t0 <= lim)) t0 = kMiss;
  if (!(t1 <= lim)) t1 = kMiss;
  if (!(t2 <= lim)) t2 = kMiss;
  if (!(t3 <= lim)) t3 = kMiss;
  // ascending by entry distance (5-exchange network)
  sort2(t0, c0, t1, c1);
  sort2(t2, c2, t3, c3);
  sort2(t0, c0, t2, c2);
  sort2(t1, c1, t3, c3);
  sort2(t1, c1, t2, c2);
  int sp = T.sp;
  if (t0 < kHitLim) {
    // push the farther ones far-to-near, descend the nearest; while every lane has room for
    // three more entries in the LDS part, the pushes go there without the overflow checks
    if (Stack::kLds >= 3 && !__any(sp > Stack::kLds - 3)) {
      if (t3 < kHitLim) stk.lds_put(sp++, c3, t3);
      if (t2 < kHitLim) stk.lds_put(sp++, c2, t2);
      if (t1 < kHitLim) stk.lds_put(sp++, c1, t1);
    } else {
      if (t3 < kHitLim) push_far(b, stk, sp, c3, t3);
      if (t2 < kHitLim) push_far(b, stk, sp, c2, t2);
      if (t1 < kHitLim) push_far(b, stk, sp, c1, t1);
    }
    T.sp = sp;
    T.node = c0;
    return false;
  }
  // pop, dropping entries that start beyond the best hit; while every lane's stack top is in
  // LDS (nearly always) the reads are LDS-only, otherwise they may reach the scratch part
  if (Stack::kLds > 0 && !__any(sp > Stack::kLds)) {
    while (sp > 0 && !(stk.lds_dist(sp - 1) <= lim)) --sp;
  } else {
    while (sp > 0 && !(stk.dist(sp - 1) <= lim)) --sp;
  }
  if (sp == 0) {
    T.sp = 0;
    T.node = kTravDone;
    return true;
  }
  --sp;
  if (Stack::kLds > 0 && sp < Stack::kLds)
    T.node = stk.lds_node(sp);
  else
    T.node = stk.node(sp);
  T.sp = sp;
  return false;
}

// Closest hit with 0 < t < tmax, on the caller's stack.
// (r from ray_pre; a caller that only moves the origin of a ray, as a periodic wrap does, can
// keep r and overwrite r.o.)
template <class Stack>
__device__ __forceinline__ Hit bvh_closest_hit(const BvhView& b, const RayPre& r, float tmax,
                                               Stack& stk) {
  Trav T;
  trav_begin(b, T, r, tmax);
  while (!trav_step(b, r, T, stk)) {
  }
  return T.h;
}

// Closest hit with 0 < t < tmax.
__device__ __forceinline__ Hit bvh_closest_hit(const BvhView& b, Vec3 o, Vec3 d, float tmax) {
  ScratchStack stk;
  return bvh_closest_hit(b, ray_pre(o, d), tmax, stk);
}

// Ericson, Real-Time Collision Detection 5.1.5: closest point of triangle abc to p.
__device__ __forceinline__ Vec3 closest_on_tri(Vec3 p, Vec3 a, Vec3 b, Vec3 c) {
  Vec3 ab = b - a, ac = c - a, ap = p - a;
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.f && d2 <= 0.f) return a;
  Vec3 bp = p - b;
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.f && d4 <= d3) return b;
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
    float v = d1 / (d1 - d3);
    return a + v * ab;
  }
  Vec3 cp = p - c;
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.f && d5 <= d6) return c;
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
    float w = d2 / (d2 - d6);
    return a + w * ac;
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return b + w * (c - b);
  }
  float den = 1.f / (va + vb + vc);
  float v = vb * den, w = vc * den;
  return a + v * ab + w * ac;
}

__device__ __forceinline__ float box_dist2(Vec3 p, float lx, float ly, float lz, float hx,
                                           float hy, float hz) {
  float dx = fmaxf(fmaxf(lx - p.x, 0.f), p.x - hx);
  float dy = fmaxf(fmaxf(ly - p.y, 0.f), p.y - hy);
  float dz = fmaxf(fmaxf(lz - p.z, 0.f), p.z - hz);
  return dx * dx + dy * dy + dz * dz;
}

struct PointHit {
  int face;
  float d2;
};

// Nearest-face ties. A grid point whose closest surface point is a vertex or an edge shared by
// several faces is equally near all of them, and which one a traversal meets first depends on
// the tree. The query therefore returns, among the faces whose distance is within
// tie_band(p) of the minimum, the one with the lowest face id, so every implementation of it
// (this one, the oracle's BVH and brute force) assigns the same face. The band, 1e-6 of the
// point's coordinate magnitude, is a few dozen float ulps of the coordinates: wide enough that
// float rounding of the distances to a shared vertex or edge stays inside it, far below any
// geometric difference.
__device__ __forceinline__ float tie_band(Vec3 p) {
  return 1.0e-6f * (fabsf(p.x) + fabsf(p.y) + fabsf(p.z) + 1.f);
}

__device__ __forceinline__ float point_face_d2(const BvhView& b, Vec3 p, int slot, int& face) {
  float4 q0 = b.tris[3 * slot], q1 = b.tris[3 * slot + 1], q2 = b.tris[3 * slot + 2];
  Vec3 c = closest_on_tri(p, v3(q0.x, q0.y, q0.z), v3(q1.x, q1.y, q1.z), v3(q2.x, q2.y, q2.z));
  Vec3 e = p - c;
  face = __float_as_int(q0.w);
  return dot(e, e);
}

// One traversal of the binary tree. Pass 0 (lim < 0): minimum squared distance below h.d2,
// in h. Pass 1: lowest face id with squared distance <= lim, in h (h.d2 its distance).
__device__ __forceinline__ void point_pass(const BvhView& b, Vec3 p, float lim, PointHit& h) {
  const bool ties = lim >= 0.f;
  int stack[kBvhStack];
  int sp = 0;
  int node = 0;
  const int max_steps = 2 * b.nf + 64;
  for (int steps = 0;; ++steps) {
    if (steps > max_steps || node < 0 || node >= b.nf - 1) {
      if (b.err) *b.err = 2;
      break;
    }
    const float cut = ties ? lim : h.d2;  // boxes farther than this cannot help
    const BvhNode* N = b.nodes + node;
    float4 A = N->a, B = N->b, C = N->c;
    int4 D = N->d;
    float dl = box_dist2(p, A.x, A.y, A.z, A.w, B.x, B.y);
    float dr = box_dist2(p, B.z, B.w, C.x, C.y, C.z, C.w);
    bool hl = ties ? dl <= cut : dl < cut, hr = ties ? dr <= cut : dr < cut;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int c = side ? D.y : D.x;
      bool& hit = side ? hr : hl;
      if (!hit || c >= 0) continue;
      hit = false;
      const int slot = ~c;
      if (slot < 0 || slot >= b.nf) continue;
      int face;
      const float d2 = point_face_d2(b, p, slot, face);
      if (ties) {
        if (d2 <= lim && (h.face < 0 || face < h.face)) {
          h.face = face;
          h.d2 = d2;
        }
      } else if (d2 < h.d2) {
        h.face = face;
        h.d2 = d2;
      }
    }
    if (!ties) {  // the other child may have become irrelevant after a leaf test
      if (hl && !(dl < h.d2)) hl = false;
      if (hr && !(dr < h.d2)) hr = false;
    }
    if (hl && hr) {
      int nearc = (dl <= dr) ? D.x : D.y;
      int farc = (dl <= dr) ? D.y : D.x;
      if (sp < kBvhStack) {
        stack[sp++] = farc;
      } else if (b.err) {
        *b.err = 1;
      }
      node = nearc;
    } else if (hl) {
      node = D.x;
    } else if (hr) {
      node = D.y;
    } else {
      if (sp == 0) break;
      node = stack[--sp];
    }
  }
}

// Squared tie limit of a minimum squared distance m: faces within tie_band(p) of sqrt(m).
__device__ __forceinline__ float tie_lim(float m, float tb, float maxd2) {
  const float r = sqrtf(m) + tb;
  return fminf(r * r, maxd2);
}

// The two passes of the query in one traversal: the running minimum m of the squared distances
// below max_dist^2, and the faces with d2 <= tie_lim(m) in kTieSlots slots. Boxes are visited
// while their distance is <= tie_lim(m); a face enters while d2 <= tie_lim(m) and its slot is
// reused once d2 > tie_lim(m) for a smaller m. Every face within tie_lim(final m) is in a box
// visited at a larger limit and keeps its slot, so the slots end holding exactly the faces the
// second pass would select. Returns false if a face found no slot (the caller then runs the
// second pass).
constexpr int kTieSlots = 8;

__device__ __forceinline__ bool point_pass_ties(const BvhView& b, Vec3 p, float maxd2, float tb,
                                                float& m, int (&cf)[kTieSlots],
                                                float (&cd)[kTieSlots]) {
  bool ok = true;
  float lim = maxd2;
  int stack[kBvhStack];
  int sp = 0;
  int node = 0;
  const int max_steps = 2 * b.nf + 64;
  for (int steps = 0;; ++steps) {
    if (steps > max_steps || node < 0 || node >= b.nf - 1) {
      if (b.err) *b.err = 2;
      break;
    }
    const BvhNode* N = b.nodes + node;
    float4 A = N->a, B = N->b, C = N->c;
    int4 D = N->d;
    float dl = box_dist2(p, A.x, A.y, A.z, A.w, B.x, B.y);
    float dr = box_dist2(p, B.z, B.w, C.x, C.y, C.z, C.w);
    bool hl = dl <= lim, hr = dr <= lim;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int c = side ? D.y : D.x;
      bool& hit = side ? hr : hl;
      if (!hit || c >= 0) continue;
      hit = false;
      const int slot = ~c;
      if (slot < 0 || slot >= b.nf) continue;
      int face;
      const float d2 = point_face_d2(b, p, slot, face);
      if (d2 < m) {
        m = d2;
        lim = tie_lim(m, tb, maxd2);
      }
      if (d2 <= lim) {
        bool placed = false;
#pragma unroll
        for (int j = 0; j < kTieSlots; ++j) {
          if (!placed && !(cd[j] <= lim)) {
            cf[j] = face;
            cd[j] = d2;
            placed = true;
          }
        }
        ok = ok && placed;
      }
    }
    if (hl && !(dl <= lim)) hl = false;
    if (hr && !(dr <= lim)) hr = false;
    if (hl && hr) {
      int nearc = (dl <= dr) ? D.x : D.y;
      int farc = (dl <= dr) ? D.y : D.x;
      if (sp < kBvhStack) {
        stack[sp++] = farc;
      } else if (b.err) {
        *b.err = 1;
      }
      node = nearc;
    } else if (hl) {
      node = D.x;
    } else if (hr) {
      node = D.y;
    } else {
      if (sp == 0) break;
      node = stack[--sp];
    }
  }
  return ok;
}

// Closest face within max_dist (distance squared < max_dist^2), ties to the lowest face id.
__device__ __forceinline__ PointHit bvh_closest_point(const BvhView& b, Vec3 p, float max_dist) {
  const float maxd2 = max_dist * max_dist;
  PointHit h{-1, maxd2};
  if (b.nf <= 0) return h;
  if (b.nf == 1) {
    int face;
    const float d2 = point_face_d2(b, p, 0, face);
    if (d2 < h.d2) h = PointHit{face, d2};
    return h;
  }
  const float tb = tie_band(p);
  float m = maxd2;
  int cf[kTieSlots];
  float cd[kTieSlots];
#pragma unroll
  for (int j = 0; j < kTieSlots; ++j) {
    cf[j] = -1;
    cd[j] = INFINITY;
  }
  const bool ok = point_pass_ties(b, p, maxd2, tb, m, cf, cd);
  if (!(m < maxd2)) return h;                 // no face below max_dist
  const float lim = tie_lim(m, tb, maxd2);
  PointHit t{-1, m};
  if (ok) {
#pragma unroll
    for (int j = 0; j < kTieSlots; ++j) {
      if (cd[j] <= lim && (t.face < 0 || cf[j] < t.face)) {
        t.face = cf[j];
        t.d2 = cd[j];
      }
    }
  } else {
    point_pass(b, p, lim, t);
  }
  return t;
}

}  // namespace petch
