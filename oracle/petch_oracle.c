/* petch_oracle.c — CPU restatement of the petch etch-loop hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline through oracle/native.py; never linked into or called by the product library.
 *
 * Contents (each section cites the reference lines it restates):
 *   Philox-4x32-10 counter RNG (Salmon et al., SC'11) and the uniform / normal draws used by
 *     every stochastic kernel here (the reference uses Warp's wp.rand_init/randf/randn streams,
 *     threed.py:361,448,520,560,616 — those streams cannot be reproduced; parity is statistical).
 *   Marching cubes with the loop-tracing rule documented in oc_cell() (replaces skimage /
 *     Warp MC, threed.py:62-100; their own triangulations are unpinned).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* =============================================================== Philox-4x32-10 ========== */
static void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0;
    uint32_t h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
    uint32_t n0 = h1 ^ c1 ^ k0, n2 = h0 ^ c3 ^ k1;
    c0 = n0; c1 = l1; c2 = n2; c3 = l0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  philox4x32_10(ctr, key, out);
}

/* =============================================================== marching cubes ========== */
/* Corner c = cx + 2cy + 4cz. Cube edges, numbered axis-major (x edges 0-3, y 4-7, z 8-11), each
 * as its (low, high) corner pair. */
static const int OC_EDGE[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                                   {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
/* The six faces, corners listed counter-clockwise seen from outside the cube:
 * x=0, x=1, y=0, y=1, z=0, z=1. */
static const int OC_FACE[6][4] = {{0, 4, 6, 2}, {1, 3, 7, 5}, {0, 1, 5, 4},
                                  {2, 6, 7, 3}, {0, 2, 3, 1}, {4, 5, 7, 6}};

static int oc_edge_id(int a, int b) {
  int lo = a < b ? a : b, hi = a < b ? b : a;
  for (int e = 0; e < 12; ++e)
    if (OC_EDGE[e][0] == lo && OC_EDGE[e][1] == hi) return e;
  return -1;
}

/* The two cube faces an edge lies on, as bits (2*axis + side) of a 6-bit mask. */
static int oc_edge_faces(int e) {
  int lo = OC_EDGE[e][0], hi = OC_EDGE[e][1], m = 0;
  for (int b = 0; b < 3; ++b)
    if (((lo >> b) & 1) == ((hi >> b) & 1)) m |= 1 << (2 * b + ((lo >> b) & 1));
  return m;
}

/* Triangulates one loop by ear clipping: repeatedly cut the first ear (L[i-1], L[i], L[i+1]),
 * in loop order, whose new diagonal L[i-1]-L[i+1] does not join two points of one cube face (such
 * a diagonal lies in the face plane and would also be cut by the neighbouring cube, giving an edge
 * with four faces); if every ear has one, cut the first. Triangles are written reversed,
 * (L[i-1], L[i+1], L[i]), so their normals point into the gas. Returns len - 2. */
static int oc_clip_loop(const int* L0, int len, int (*tri)[3]) {
  int L[12], nt = 0;
  for (int i = 0; i < len; ++i) L[i] = L0[i];
  while (len > 3) {
    int pick = 0;
    for (int i = 0; i < len; ++i) {
      int a = L[(i + len - 1) % len], b = L[(i + 1) % len];
      if (!(oc_edge_faces(a) & oc_edge_faces(b))) {
        pick = i;
        break;
      }
    }
    tri[nt][0] = L[(pick + len - 1) % len];
    tri[nt][1] = L[(pick + 1) % len];
    tri[nt][2] = L[pick];
    ++nt;
    for (int i = pick; i + 1 < len; ++i) L[i] = L[i + 1];
    --len;
  }
  if (len == 3) {
    tri[nt][0] = L[0];
    tri[nt][1] = L[2];
    tri[nt][2] = L[1];
    ++nt;
  }
  return nt;
}

/* Surface of one cube as triangles of cube-edge ids (returns the count, <= 10).
 * A corner is solid iff f > 0. On each face, walking the corner cycle counter-clockwise seen from
 * outside, the surface segment starts where the walk leaves a solid corner and ends where it
 * enters one (solid on the left). With two solid corners on one diagonal (ambiguous face), the
 * solid corners are joined through the face iff f(solid1)*f(solid2) > f(gas1)*f(gas2): then a
 * segment starting on the edge after a solid corner ends on the next edge (around the gas
 * corner between them); otherwise it ends on the edge before that solid corner. Segments chain
 * into loops; loops are walked from their smallest edge id and triangulated by oc_clip_loop. */
static int oc_cell(const float f[8], int tri[10][3]) {
  int solid[8], ns = 0;
  for (int c = 0; c < 8; ++c) {
    solid[c] = f[c] > 0.0f;
    ns += solid[c];
  }
  if (ns == 0 || ns == 8) return 0;
  int next[12];
  for (int e = 0; e < 12; ++e) next[e] = -1;
  for (int fc = 0; fc < 6; ++fc) {
    const int* cy = OC_FACE[fc];
    int ex[4], cross = 0;
    for (int k = 0; k < 4; ++k) {
      ex[k] = oc_edge_id(cy[k], cy[(k + 1) % 4]);
      cross += solid[cy[k]] != solid[cy[(k + 1) % 4]];
    }
    if (cross == 2) {
      int from = -1, to = -1;
      for (int k = 0; k < 4; ++k) {
        int a = solid[cy[k]], b = solid[cy[(k + 1) % 4]];
        if (a && !b) from = ex[k];
        if (!a && b) to = ex[k];
      }
      next[from] = to;
    } else if (cross == 4) {
      int s0 = solid[cy[0]] ? 0 : 1; /* position of a solid corner in the cycle */
      float ps = f[cy[s0]] * f[cy[s0 + 2]];
      float pg = f[cy[1 - s0]] * f[cy[3 - s0]];
      int join = ps > pg;
      for (int k = 0; k < 4; ++k) {
        if (solid[cy[k]] && !solid[cy[(k + 1) % 4]])
          next[ex[k]] = join ? ex[(k + 1) % 4] : ex[(k + 3) % 4];
      }
    }
  }
  int used[12] = {0}, nt = 0;
  for (int e0 = 0; e0 < 12; ++e0) {
    if (next[e0] < 0 || used[e0]) continue;
    int L[12], len = 0, e = e0;
    do {
      L[len++] = e;
      used[e] = 1;
      e = next[e];
    } while (e != e0 && e >= 0 && len < 12);
    nt += oc_clip_loop(L, len, tri + nt);
  }
  return nt;
}

/* Whole-grid marching cubes. Vertex ids follow (node, axis) order of the grid edge they lie on,
 * faces follow cell order. Positions: ((double)index + t) * dx with t = p0/(p0-p1), rounded to
 * float once. With cap_v/cap_f too small only the counts are returned (rc 1). */
int oracle_marching_cubes(const float* phi, int nx, int ny, int nz, double dx, float* verts,
                          int* faces, long long cap_v, long long cap_f, long long* nv_out,
                          long long* nf_out) {
  *nv_out = 0;
  *nf_out = 0;
  if (nx < 2 || ny < 2 || nz < 2) return 0;
  long long N = (long long)nx * ny * nz, sx = (long long)ny * nz, sy = nz;
  int* vid = (int*)malloc(sizeof(int) * 3 * N);
  if (!vid) return -1;
  long long nv = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        long long n = i * sx + j * sy + k;
        int idx[3] = {i, j, k}, dim[3] = {nx, ny, nz};
        long long st[3] = {sx, sy, 1};
        for (int a = 0; a < 3; ++a) {
          vid[3 * n + a] = -1;
          if (idx[a] >= dim[a] - 1) continue;
          float p0 = phi[n], p1 = phi[n + st[a]];
          if ((p0 > 0.0f) == (p1 > 0.0f)) continue;
          if (nv < cap_v && verts) {
            double t = (double)p0 / ((double)p0 - (double)p1);
            double c[3] = {(double)i, (double)j, (double)k};
            c[a] = c[a] + t;
            for (int d = 0; d < 3; ++d) verts[3 * nv + d] = (float)(c[d] * dx);
          }
          vid[3 * n + a] = (int)nv++;
        }
      }
  long long nf = 0;
  for (int i = 0; i + 1 < nx; ++i)
    for (int j = 0; j + 1 < ny; ++j)
      for (int k = 0; k + 1 < nz; ++k) {
        long long n = i * sx + j * sy + k;
        float f[8];
        for (int c = 0; c < 8; ++c)
          f[c] = phi[n + (c & 1) * sx + ((c >> 1) & 1) * sy + ((c >> 2) & 1)];
        int tri[10][3];
        int nt = oc_cell(f, tri);
        for (int q = 0; q < nt; ++q) {
          if (nf < cap_f && faces) {
            for (int m = 0; m < 3; ++m) {
              int e = tri[q][m];
              int lo = OC_EDGE[e][0];
              int a = e / 4;
              long long base = n + (lo & 1) * sx + ((lo >> 1) & 1) * sy + ((lo >> 2) & 1);
              faces[3 * nf + m] = vid[3 * base + a];
            }
          }
          ++nf;
        }
      }
  free(vid);
  *nv_out = nv;
  *nf_out = nf;
  return (nv > cap_v || nf > cap_f) ? 1 : 0;
}

/* Cube-local triangulation for one corner configuration (used by the table tests). */
int oracle_mc_cell(const float* f8, int* tri30) {
  int tri[10][3];
  int nt = oc_cell(f8, tri);
  for (int q = 0; q < nt; ++q)
    for (int m = 0; m < 3; ++m) tri30[3 * q + m] = tri[q][m];
  return nt;
}

/* =============================================================== brute-force mesh queries == */
/* Closest two-sided hit with 0 < t < tmax over all faces (Moller-Trumbore in double): the
 * semantics of wp.mesh_query_ray (threed.py:306,366,524,569,629,916). */
static int oc_ray_tri_d(const double o[3], const double d[3], const double a[3], const double b[3],
                        const double c[3], double* t) {
  double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2],
                 d[0] * e2[1] - d[1] * e2[0]};
  double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
  if (det == 0.0) return 0;
  double inv = 1.0 / det;
  double s[3] = {o[0] - a[0], o[1] - a[1], o[2] - a[2]};
  double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
  if (u < 0.0 || u > 1.0) return 0;
  double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2],
                 s[0] * e1[1] - s[1] * e1[0]};
  double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
  if (v < 0.0 || u + v > 1.0) return 0;
  *t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
  return *t > 0.0;
}

static void oc_vert(const float* verts, int i, double out[3]) {
  out[0] = verts[3 * i];
  out[1] = verts[3 * i + 1];
  out[2] = verts[3 * i + 2];
}

void oracle_closest_hit_brute(const float* verts, const int* faces, int nf, const float* orig,
                              const float* dir, int n, double tmax, int* face_out,
                              double* t_out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    double o[3] = {orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]};
    double d[3] = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
    int best = -1;
    double bt = tmax;
    for (int f = 0; f < nf; ++f) {
      double a[3], b[3], c[3], t;
      oc_vert(verts, faces[3 * f], a);
      oc_vert(verts, faces[3 * f + 1], b);
      oc_vert(verts, faces[3 * f + 2], c);
      if (oc_ray_tri_d(o, d, a, b, c, &t) && t < bt) {
        bt = t;
        best = f;
      }
    }
    face_out[i] = best;
    t_out[i] = best >= 0 ? bt : -1.0;
  }
}

/* Closest point of triangle abc to p (Ericson, Real-Time Collision Detection 5.1.5), double. */
static void oc_closest_on_tri(const double p[3], const double a[3], const double b[3],
                              const double c[3], double out[3]) {
#define OC_DOT(x, y) ((x)[0] * (y)[0] + (x)[1] * (y)[1] + (x)[2] * (y)[2])
  double ab[3], ac[3], ap[3], bp[3], cp[3];
  for (int k = 0; k < 3; ++k) {
    ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; ap[k] = p[k] - a[k];
    bp[k] = p[k] - b[k]; cp[k] = p[k] - c[k];
  }
  double d1 = OC_DOT(ab, ap), d2 = OC_DOT(ac, ap);
  if (d1 <= 0 && d2 <= 0) { memcpy(out, a, 24); return; }
  double d3 = OC_DOT(ab, bp), d4 = OC_DOT(ac, bp);
  if (d3 >= 0 && d4 <= d3) { memcpy(out, b, 24); return; }
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    double v = d1 / (d1 - d3);
    for (int k = 0; k < 3; ++k) out[k] = a[k] + v * ab[k];
    return;
  }
  double d5 = OC_DOT(ab, cp), d6 = OC_DOT(ac, cp);
  if (d6 >= 0 && d5 <= d6) { memcpy(out, c, 24); return; }
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) {
    double w = d2 / (d2 - d6);
    for (int k = 0; k < 3; ++k) out[k] = a[k] + w * ac[k];
    return;
  }
  double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int k = 0; k < 3; ++k) out[k] = b[k] + w * (c[k] - b[k]);
    return;
  }
  double den = 1.0 / (va + vb + vc);
  double v = vb * den, w = vc * den;
  for (int k = 0; k < 3; ++k) out[k] = a[k] + v * ab[k] + w * ac[k];
#undef OC_DOT
}

/* Closest face to each point (wp.mesh_query_point_no_sign, threed.py:1328,1626); ties keep the
 * lowest face id. */
/* Nearest-face ties: among the faces whose distance is within the tie band of the minimum, the
 * lowest face id (a point nearest to a shared vertex or edge is equally near several faces; the
 * rule makes the choice independent of the search order). The band is 1e-6 of the point's
 * coordinate magnitude, as on the device. Returns the squared-distance limit of the band. */
static double oc_tie_lim(const double p[3], double d2min) {
  const double band = 1.0e-6 * (fabs(p[0]) + fabs(p[1]) + fabs(p[2]) + 1.0);
  const double r = sqrt(d2min) + band;
  return r * r;
}

void oracle_closest_point_brute(const float* verts, const int* faces, int nf, const float* pts,
                                int n, int* face_out, double* dist_out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    double bd = 1e300;
    double* dd = (double*)malloc(sizeof(double) * (nf > 0 ? nf : 1));
    for (int f = 0; f < nf; ++f) {
      double a[3], b[3], c[3], q[3];
      oc_vert(verts, faces[3 * f], a);
      oc_vert(verts, faces[3 * f + 1], b);
      oc_vert(verts, faces[3 * f + 2], c);
      oc_closest_on_tri(p, a, b, c, q);
      dd[f] = (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) +
              (p[2] - q[2]) * (p[2] - q[2]);
      if (dd[f] < bd) bd = dd[f];
    }
    int best = -1;
    const double lim = oc_tie_lim(p, bd);
    for (int f = 0; f < nf && best < 0; ++f)
      if (dd[f] <= lim) best = f;
    face_out[i] = best;
    dist_out[i] = best >= 0 ? sqrt(dd[best]) : -1.0;
    free(dd);
  }
}

/* =============================================================== per-ray RNG stream ======== */
/* Same stream layout as the device: key = seed, counter = (block, ray, 0, 0), words in order;
 * a Russian-roulette neutral's k-th bounce takes the words of block (k, ray, 1, 0). */
typedef struct {
  uint32_t key[2], ray, blk, w[4];
  int pos;
} oc_rng;

static void oc_rng_init(oc_rng* r, uint64_t seed, uint32_t ray) {
  r->key[0] = (uint32_t)seed;
  r->key[1] = (uint32_t)(seed >> 32);
  r->ray = ray;
  r->blk = 0;
  r->pos = 4;
}
static uint32_t oc_next(oc_rng* r) {
  if (r->pos == 4) {
    uint32_t c[4] = {r->blk++, r->ray, 0u, 0u};
    philox4x32_10(c, r->key, r->w);
    r->pos = 0;
  }
  return r->w[r->pos++];
}
static float oc_u24(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-8f; }
static float oc_uniform(oc_rng* r) { return oc_u24(oc_next(r)); }
/* the four words of block (blk, ray, lane, 0): a neutral's k-th bounce draws from (k, ray, 1) */
static void oc_block(uint64_t seed, uint32_t ray, uint32_t blk, uint32_t lane, uint32_t out[4]) {
  uint32_t c[4] = {blk, ray, lane, 0u}, key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  philox4x32_10(c, key, out);
}
static float oc_normal(oc_rng* r) {
  float u1 = (float)((oc_next(r) >> 8) + 1u) * 5.9604644775390625e-8f;
  float u2 = oc_uniform(r);
  return sqrtf(-2.f * logf(u1)) * cosf(6.28318530717958648f * u2);
}

/* =============================================================== CPU BVH (oracle) ========== */
/* Top-down median split over centroids, <= 4 faces per leaf; closest hit in double precision
 * (checked against oracle_closest_hit_brute by the tests). Independent of the device LBVH. */
typedef struct {
  float lo[3], hi[3];
  int left, right; /* children, or -1 */
  int start, count; /* leaf face range in order[] */
} oc_node;

typedef struct {
  int nf, nnodes;
  oc_node* nodes;
  int* order;
  double* tri; /* 9 doubles per face: v0 v1 v2 */
} oc_bvh;

static float* oc_cen_tmp;
static int oc_cmp_axis;
static int oc_cmp(const void* a, const void* b) {
  float ca = oc_cen_tmp[3 * (*(const int*)a) + oc_cmp_axis];
  float cb = oc_cen_tmp[3 * (*(const int*)b) + oc_cmp_axis];
  if (ca < cb) return -1;
  if (ca > cb) return 1;
  return (*(const int*)a) - (*(const int*)b);
}

static int oc_build_rec(oc_bvh* B, float* cen, int start, int count) {
  int id = B->nnodes++;
  oc_node* nd = &B->nodes[id];
  for (int k = 0; k < 3; ++k) { nd->lo[k] = 3e38f; nd->hi[k] = -3e38f; }
  for (int i = start; i < start + count; ++i) {
    const double* t = B->tri + 9 * B->order[i];
    for (int v = 0; v < 3; ++v)
      for (int k = 0; k < 3; ++k) {
        float x = (float)t[3 * v + k];
        if (x < nd->lo[k]) nd->lo[k] = x;
        if (x > nd->hi[k]) nd->hi[k] = x;
      }
  }
  nd->left = nd->right = -1;
  nd->start = start;
  nd->count = count;
  if (count <= 4) return id;
  float clo[3] = {3e38f, 3e38f, 3e38f}, chi[3] = {-3e38f, -3e38f, -3e38f};
  for (int i = start; i < start + count; ++i)
    for (int k = 0; k < 3; ++k) {
      float c = cen[3 * B->order[i] + k];
      if (c < clo[k]) clo[k] = c;
      if (c > chi[k]) chi[k] = c;
    }
  int ax = 0;
  for (int k = 1; k < 3; ++k)
    if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
  oc_cen_tmp = cen;
  oc_cmp_axis = ax;
  qsort(B->order + start, count, sizeof(int), oc_cmp);
  int half = count / 2;
  int l = oc_build_rec(B, cen, start, half);
  int r = oc_build_rec(B, cen, start + half, count - half);
  B->nodes[id].left = l;
  B->nodes[id].right = r;
  return id;
}

void* oracle_bvh_build(const float* verts, const int* faces, int nf) {
  oc_bvh* B = (oc_bvh*)calloc(1, sizeof(oc_bvh));
  B->nf = nf;
  B->nodes = (oc_node*)calloc(nf > 0 ? 2 * nf : 1, sizeof(oc_node));
  B->order = (int*)malloc(sizeof(int) * (nf > 0 ? nf : 1));
  B->tri = (double*)malloc(sizeof(double) * 9 * (nf > 0 ? nf : 1));
  float* cen = (float*)malloc(sizeof(float) * 3 * (nf > 0 ? nf : 1));
  for (int f = 0; f < nf; ++f) {
    B->order[f] = f;
    for (int v = 0; v < 3; ++v)
      for (int k = 0; k < 3; ++k) B->tri[9 * f + 3 * v + k] = verts[3 * faces[3 * f + v] + k];
    for (int k = 0; k < 3; ++k)
      cen[3 * f + k] = (float)((B->tri[9 * f + k] + B->tri[9 * f + 3 + k] + B->tri[9 * f + 6 + k]) / 3.0);
  }
  if (nf > 0) oc_build_rec(B, cen, 0, nf);
  free(cen);
  return B;
}

void oracle_bvh_free(void* h) {
  oc_bvh* B = (oc_bvh*)h;
  if (!B) return;
  free(B->nodes);
  free(B->order);
  free(B->tri);
  free(B);
}

static int oc_box_hit(const oc_node* nd, const double o[3], const double inv[3], double tbest) {
  double t0 = 0.0, t1 = tbest;
  for (int k = 0; k < 3; ++k) {
    double a = ((double)nd->lo[k] - o[k]) * inv[k], b = ((double)nd->hi[k] - o[k]) * inv[k];
    if (a > b) { double t = a; a = b; b = t; }
    if (a > t0) t0 = a;
    if (b < t1) t1 = b;
  }
  return t0 <= t1 * (1.0 + 1e-9) + 1e-12;
}

/* Closest hit 0 < t < tmax; returns the face (or -1) and sets *t. */
static int oc_closest(const oc_bvh* B, const double o[3], const double d[3], double tmax,
                      double* tout) {
  int best = -1;
  double bt = tmax;
  if (B->nf == 0) return -1;
  double inv[3];
  for (int k = 0; k < 3; ++k) inv[k] = 1.0 / (fabs(d[k]) > 1e-300 ? d[k] : 1e-300);
  int stack[128], sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const oc_node* nd = &B->nodes[stack[--sp]];
    if (!oc_box_hit(nd, o, inv, bt)) continue;
    if (nd->left < 0) {
      for (int i = nd->start; i < nd->start + nd->count; ++i) {
        int f = B->order[i];
        const double* t = B->tri + 9 * f;
        double tt;
        if (oc_ray_tri_d(o, d, t, t + 3, t + 6, &tt) && tt < bt) {
          bt = tt;
          best = f;
        }
      }
    } else {
      stack[sp++] = nd->left;
      stack[sp++] = nd->right;
    }
  }
  *tout = bt;
  return best;
}

void oracle_bvh_closest_hit(void* h, const float* orig, const float* dir, int n, double tmax,
                            int* face_out, double* t_out) {
  const oc_bvh* B = (const oc_bvh*)h;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    double o[3] = {orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]};
    double d[3] = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
    double t;
    int f = oc_closest(B, o, d, tmax, &t);
    face_out[i] = f;
    t_out[i] = f >= 0 ? t : -1.0;
  }
}

/* =============================================================== traces ==================== */
typedef struct { float Lx, Ly, Lz, z_src; int periodic; } oc_box;
typedef struct { int kind; float sigma, exponent; } oc_source;
typedef struct {
  float meanE, sigmaE, A_sp, Eth_sp, B_sp, A_ie, Eth_ie, A_p, Eth_p;
  float inflect, minang, n_l, A_en, minE, thrmin, thrmax;
} oc_ion;

typedef struct { float x, y, z; } f3;
static f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static f3 add3(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static f3 sub3(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static f3 scl(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static f3 crs(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static f3 nrmz(f3 a) {
  float l = sqrtf(dot3(a, a));
  return l > 0.f ? scl(1.f / l, a) : mk(0.f, 0.f, 0.f);
}

/* Source ray: threed.py:449-458 (Gaussian ion / cosine neutral) and the power-cosine option. */
static void oc_source_ray(oc_rng* r, const oc_box* b, const oc_source* s, f3* o, f3* d) {
  float ox = oc_uniform(r) * b->Lx;
  float oy = oc_uniform(r) * b->Ly;
  *o = mk(ox, oy, b->z_src);
  if (s->kind == 1) {
    float ax = oc_normal(r) * s->sigma;
    float ay = oc_normal(r) * s->sigma;
    *d = nrmz(mk(sinf(ax), sinf(ay), -cosf(ax) * cosf(ay)));
  } else if (s->kind == 2) {
    float L = logf(fmaxf(oc_uniform(r), 1.0e-30f)) / (s->exponent + 1.f);
    float ct = expf(L);
    float st = sqrtf(fmaxf(-expm1f(L) * (1.f + ct), 0.f));
    float ph = 6.2831853f * oc_uniform(r);
    *d = nrmz(mk(st * cosf(ph), st * sinf(ph), -ct));
  } else {
    float r1 = oc_uniform(r), r2 = oc_uniform(r);
    float ct = sqrtf(r1), st = sqrtf(1.f - r1), ph = 6.2831853f * r2;
    *d = nrmz(mk(st * cosf(ph), st * sinf(ph), -ct));
  }
}

void oracle_gen_source(const oc_box* b, const oc_source* s, int n, uint64_t seed, float* origin,
                       float* dir) {
  for (int p = 0; p < n; ++p) {
    oc_rng r;
    oc_rng_init(&r, seed, (uint32_t)p);
    f3 o, d;
    oc_source_ray(&r, b, s, &o, &d);
    origin[3 * p] = o.x; origin[3 * p + 1] = o.y; origin[3 * p + 2] = o.z;
    dir[3 * p] = d.x; dir[3 * p + 1] = d.y; dir[3 * p + 2] = d.z;
  }
}

/* Closest hit from float origin/direction; returns face, t and the unit face normal. */
static int oc_query(const oc_bvh* B, f3 o, f3 d, float* t, f3* n) {
  double od[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z}, tt;
  int f = oc_closest(B, od, dd, 1.0e6, &tt);
  if (f < 0) return -1;
  *t = (float)tt;
  const double* q = B->tri + 9 * f;
  f3 a = mk((float)q[0], (float)q[1], (float)q[2]);
  f3 b = mk((float)q[3], (float)q[4], (float)q[5]);
  f3 c = mk((float)q[6], (float)q[7], (float)q[8]);
  *n = nrmz(crs(sub3(b, a), sub3(c, a)));
  return f;
}

/* _apply_bc (threed.py:289-349) followed by the caller's mesh_query_ray. */
static int oc_query_bc(const oc_bvh* B, const oc_box* b, f3* o, f3 d, float* t, f3* n) {
  if (!b->periodic) return oc_query(B, *o, d, t, n);
  for (int w = 0; w < 256; ++w) {
    int f = oc_query(B, *o, d, t, n);
    if (f >= 0) return f;
    float tx = 1.0e30f, ty = 1.0e30f, tz = 1.0e30f;
    if (d.x > 1.0e-9f) tx = (b->Lx - o->x) / d.x;
    if (d.x < -1.0e-9f) tx = (0.f - o->x) / d.x;
    if (d.y > 1.0e-9f) ty = (b->Ly - o->y) / d.y;
    if (d.y < -1.0e-9f) ty = (0.f - o->y) / d.y;
    if (d.z > 1.0e-9f) tz = (b->Lz - o->z) / d.z;
    if (tx < tz && tx < ty && tx > 1.0e-7f) {
      float hy = o->y + tx * d.y, hz = o->z + tx * d.z;
      *o = mk(d.x < 0.f ? b->Lx - 1.0e-4f : 1.0e-4f, hy, hz);
    } else if (ty < tz && ty < 1.0e29f && ty > 1.0e-7f) {
      float hx = o->x + ty * d.x, hz = o->z + ty * d.z;
      *o = mk(hx, d.y < 0.f ? b->Ly - 1.0e-4f : 1.0e-4f, hz);
    } else {
      return -1;
    }
  }
  return oc_query(B, *o, d, t, n);
}

/* cosine re-emission about n (threed.py:395-405) from the uniforms r1, r2 */
static f3 oc_cosine_u(f3 n, float r1, float r2) {
  f3 a = fabsf(n.x) > 0.9f ? mk(0.f, 1.f, 0.f) : mk(1.f, 0.f, 0.f);
  f3 t = nrmz(crs(a, n));
  f3 b = crs(n, t);
  float ct = sqrtf(r1), st = sqrtf(1.f - r1), ph = 6.2831853f * r2;
  return add3(add3(scl(st * cosf(ph), t), scl(st * sinf(ph), b)), scl(ct, n));
}
static f3 oc_cosine(oc_rng* r, f3 n) {
  float r1 = oc_uniform(r), r2 = oc_uniform(r);
  return oc_cosine_u(n, r1, r2);
}

/* One ray of _trace3d_cov_rr (threed.py:549-596): deposits into acc (per-face, double) and
 * reports the ray's hit count and deposited weight when nhit / wsum are given. */
static void oc_rr_ray(const oc_bvh* B, const oc_box* b, const oc_source* s, const float* bare,
                      float beta, uint64_t seed, int p, double* acc, int* nhit, double* wsum) {
  oc_rng r;
  oc_rng_init(&r, seed, (uint32_t)p);
  f3 o, d, nn;
  oc_source_ray(&r, b, s, &o, &d);
  float w = 1.f, t;
  int k = 0;
  double tot = 0.0;
  for (int bounce = 0; bounce < 1024; ++bounce) {
    int f = oc_query_bc(B, b, &o, d, &t, &nn);
    if (f < 0) break;
    if (dot3(nn, d) > 0.f) nn = scl(-1.f, nn);
    acc[f] += w;
    ++k;
    tot += w;
    float S = bare[f] * beta;
    w = w - w * S;
    if (w <= 1.0e-6f) break;
    /* the bounce's words: roulette, then the lobe's two uniforms */
    uint32_t u[4];
    oc_block(seed, (uint32_t)p, (uint32_t)bounce + 1u, 1u, u);
    if (w < 0.1f) {
      if (oc_u24(u[0]) < (1.f - w / 0.3f)) break;
      w = 0.3f;
    }
    f3 hit = add3(o, scl(t, d));
    d = oc_cosine_u(nn, oc_u24(u[1]), oc_u24(u[2]));
    o = add3(hit, scl(1.0e-4f, nn));
  }
  if (nhit) *nhit = k;
  if (wsum) *wsum = tot;
}

/* _trace3d_cov_rr (threed.py:549-596); per-face sums in double. */
void oracle_trace_neutral_rr(void* h, const oc_box* b, const oc_source* s, const float* bare,
                             float beta, int n, uint64_t seed, double* fl) {
  const oc_bvh* B = (const oc_bvh*)h;
  int F = B->nf;
  memset(fl, 0, sizeof(double) * F);
#pragma omp parallel
  {
    double* acc = (double*)calloc(F > 0 ? F : 1, sizeof(double));
#pragma omp for schedule(dynamic, 256)
    for (int p = 0; p < n; ++p) oc_rr_ray(B, b, s, bare, beta, seed, p, acc, NULL, NULL);
#pragma omp critical
    for (int f = 0; f < F; ++f) fl[f] += acc[f];
    free(acc);
  }
}

/* oracle_trace_neutral_rr with each ray's hit count and deposited weight (test instrument, not
 * in the reference; the same per-ray code). */
void oracle_trace_neutral_rr_paths(void* h, const oc_box* b, const oc_source* s,
                                   const float* bare, float beta, int n, uint64_t seed,
                                   double* fl, int* nhit, double* wsum) {
  const oc_bvh* B = (const oc_bvh*)h;
  int F = B->nf;
  memset(fl, 0, sizeof(double) * F);
#pragma omp parallel
  {
    double* acc = (double*)calloc(F > 0 ? F : 1, sizeof(double));
#pragma omp for schedule(dynamic, 256)
    for (int p = 0; p < n; ++p) oc_rr_ray(B, b, s, bare, beta, seed, p, acc, nhit + p, wsum + p);
#pragma omp critical
    for (int f = 0; f < F; ++f) fl[f] += acc[f];
    free(acc);
  }
}

/* _trace3d_cov (threed.py:511-546), cov_transport = 'cap': 1 per hit, stick with probability
 * bare * beta (word 0 of the bounce's block), at most n_reemit re-emissions, no boundary
 * condition (the box's periodic flag is ignored); per-face sums in double. */
void oracle_trace_neutral_cap(void* h, const oc_box* b, const oc_source* s, const float* bare,
                              float beta, int n_reemit, int n, uint64_t seed, double* fl) {
  const oc_bvh* B = (const oc_bvh*)h;
  int F = B->nf;
  oc_box open = *b;
  open.periodic = 0;
  memset(fl, 0, sizeof(double) * (F > 0 ? F : 1));
#pragma omp parallel
  {
    double* acc = (double*)calloc(F > 0 ? F : 1, sizeof(double));
#pragma omp for schedule(dynamic, 256)
    for (int p = 0; p < n; ++p) {
      oc_rng r;
      oc_rng_init(&r, seed, (uint32_t)p);
      f3 o, d, nn;
      oc_source_ray(&r, &open, s, &o, &d);
      float t;
      for (int bounce = 0; bounce <= n_reemit; ++bounce) {
        int f = oc_query_bc(B, &open, &o, d, &t, &nn);
        if (f < 0) break;
        if (dot3(nn, d) > 0.f) nn = scl(-1.f, nn);
        acc[f] += 1.0;
        uint32_t u[4];
        oc_block(seed, (uint32_t)p, (uint32_t)bounce + 1u, 1u, u);
        if (oc_u24(u[0]) < bare[f] * beta || bounce == n_reemit) break;
        f3 hit = add3(o, scl(t, d));
        d = oc_cosine_u(nn, oc_u24(u[1]), oc_u24(u[2]));
        o = add3(hit, scl(1.0e-4f, nn));
      }
    }
#pragma omp critical
    for (int f = 0; f < F; ++f) fl[f] += acc[f];
    free(acc);
  }
}

/* Redeposition (mc_redep_3d, threed.py:1582-1605). The emission weight of face f is
 * max(rate_f area_f, 0), a non-finite product counting as 0; cum holds its running sum in order
 * (float64). Ray p draws its face from word 0 and 1 of its block 0 as a 53-bit uniform x: the
 * first f with cum[f] > x * cum[nf - 1] (the last face when none is), rng.choice's rule. */
static double oc_redep_weight(const float* rate, const float* area, int f) {
  double w = (double)rate[f] * (double)area[f];
  return (w > 0.0 && w <= 1.0e300) ? w : 0.0;
}
static int oc_redep_face(const double* cum, int nf, const uint32_t u[4]) {
  double x = ((double)(u[0] >> 5) * 67108864.0 + (double)(u[1] >> 6)) *
             (1.0 / 9007199254740992.0);
  double target = x * cum[nf - 1];
  int lo = 0, hi = nf - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (cum[mid] > target) hi = mid; else lo = mid + 1;
  }
  return lo;
}

/* The face each of n rays leaves (test instrument: the emission draw alone). */
void oracle_redep_faces(const float* rate, const float* area, int nf, int n, uint64_t seed,
                        int* face) {
  double* cum = (double*)malloc(sizeof(double) * (nf > 0 ? nf : 1));
  double c = 0.0;
  for (int f = 0; f < nf; ++f) { c += oc_redep_weight(rate, area, f); cum[f] = c; }
  for (int p = 0; p < n; ++p) {
    uint32_t u[4];
    oc_block(seed, (uint32_t)p, 0u, 0u, u);
    face[p] = nf > 0 ? oc_redep_face(cum, nf, u) : -1;
  }
  free(cum);
}

/* n product rays: from face f (drawn as above) at centroid + 1e-3 n_f along the into-gas normal
 * n_f, cosine lobe about n_f from words 2 and 3 of block 0, then the RR walk of
 * _trace3d_cov_rr with S = s_redep on every face in an open box. out = deposit * (total / n) /
 * max(area, 1e-9), all 0 when total <= 0. */
void oracle_trace_redeposition(void* h, const float* cen, const float* gn, const float* area,
                               const float* rate, float s_redep, int n, uint64_t seed,
                               double* out) {
  const oc_bvh* B = (const oc_bvh*)h;
  int F = B->nf;
  memset(out, 0, sizeof(double) * (F > 0 ? F : 1));
  if (F == 0 || n <= 0) return;
  double* cum = (double*)malloc(sizeof(double) * F);
  double c = 0.0;
  for (int f = 0; f < F; ++f) { c += oc_redep_weight(rate, area, f); cum[f] = c; }
  const double tot = cum[F - 1];
  if (tot <= 0.0) { free(cum); return; }
  const oc_box box = {1.f, 1.f, 1.f, 1.f, 0};
#pragma omp parallel
  {
    double* acc = (double*)calloc(F, sizeof(double));
#pragma omp for schedule(dynamic, 256)
    for (int p = 0; p < n; ++p) {
      uint32_t u[4];
      oc_block(seed, (uint32_t)p, 0u, 0u, u);
      int f0 = oc_redep_face(cum, F, u);
      f3 nn0 = mk(gn[3 * f0], gn[3 * f0 + 1], gn[3 * f0 + 2]);
      f3 o = add3(mk(cen[3 * f0], cen[3 * f0 + 1], cen[3 * f0 + 2]), scl(1.0e-3f, nn0));
      f3 d = oc_cosine_u(nn0, oc_u24(u[2]), oc_u24(u[3])), nn;
      float w = 1.f, t;
      for (int bounce = 0; bounce < 1024; ++bounce) {
        int f = oc_query_bc(B, &box, &o, d, &t, &nn);
        if (f < 0) break;
        if (dot3(nn, d) > 0.f) nn = scl(-1.f, nn);
        acc[f] += w;
        w = w - w * s_redep;
        if (w <= 1.0e-6f) break;
        uint32_t v[4];
        oc_block(seed, (uint32_t)p, (uint32_t)bounce + 1u, 1u, v);
        if (w < 0.1f) {
          if (oc_u24(v[0]) < (1.f - w / 0.3f)) break;
          w = 0.3f;
        }
        f3 hit = add3(o, scl(t, d));
        d = oc_cosine_u(nn, oc_u24(v[1]), oc_u24(v[2]));
        o = add3(hit, scl(1.0e-4f, nn));
      }
    }
#pragma omp critical
    for (int f = 0; f < F; ++f) out[f] += acc[f];
    free(acc);
  }
  for (int f = 0; f < F; ++f) {
    double a = (double)area[f] > 1.0e-9 ? (double)area[f] : 1.0e-9;
    out[f] = out[f] * (tot / (double)n) / a;
  }
  free(cum);
}

/* oracle_trace_neutral_rr with the decisions frozen (test instrument, not in the reference): the
 * paths, roulettes and terminations are those of the sticking bare_dec, while the deposited
 * weights follow bare_w (w_{k+1} = w_k (1 - bare_w beta); a roulette survivor's weight is
 * divided by its survival probability p = w_dec / 0.3, p taken from the bare_dec ray, which is
 * the trace's 0.3 when bare_w = bare_dec). At bare_w = bare_dec it is oracle_trace_neutral_rr;
 * its central differences in bare_w are the derivative the path adjoint computes, with no
 * decision flipping. */
void oracle_trace_neutral_rr_frozen(void* h, const oc_box* b, const oc_source* s,
                                    const float* bare_dec, const float* bare_w, float beta, int n,
                                    uint64_t seed, double* fl) {
  const oc_bvh* B = (const oc_bvh*)h;
  int F = B->nf;
  memset(fl, 0, sizeof(double) * F);
#pragma omp parallel
  {
    double* acc = (double*)calloc(F > 0 ? F : 1, sizeof(double));
#pragma omp for schedule(dynamic, 256)
    for (int p = 0; p < n; ++p) {
      oc_rng r;
      oc_rng_init(&r, seed, (uint32_t)p);
      f3 o, d, nn;
      oc_source_ray(&r, b, s, &o, &d);
      float w = 1.f, t;
      double ww = 1.0;
      for (int bounce = 0; bounce < 1024; ++bounce) {
        int f = oc_query_bc(B, b, &o, d, &t, &nn);
        if (f < 0) break;
        if (dot3(nn, d) > 0.f) nn = scl(-1.f, nn);
        acc[f] += ww;
        w = w - w * (bare_dec[f] * beta);
        ww = ww * (1.0 - (double)bare_w[f] * beta);
        if (w <= 1.0e-6f) break;
        uint32_t u[4];
        oc_block(seed, (uint32_t)p, (uint32_t)bounce + 1u, 1u, u);
        if (w < 0.1f) {
          if (oc_u24(u[0]) < (1.f - w / 0.3f)) break;
          ww = ww * (0.3 / (double)w); /* survivors carry w / p with p = w / 0.3 frozen */
          w = 0.3f;
        }
        f3 hit = add3(o, scl(t, d));
        d = oc_cosine_u(nn, oc_u24(u[1]), oc_u24(u[2]));
        o = add3(hit, scl(1.0e-4f, nn));
      }
    }
#pragma omp critical
    for (int f = 0; f < F; ++f) fl[f] += acc[f];
    free(acc);
  }
}

/* Path adjoint of oracle_trace_neutral_rr (not in the reference): gbare = d loss / d bare for
 * gfl = d loss / d fl, with every path's geometry and its roulette / termination decisions held
 * fixed and a roulette survivor's weight read as w / p with its survival probability
 * p = w / 0.3 held fixed (the value is the trace's 0.3; the derivative then carries on through
 * the roulette, and its expectation is the derivative of the expected flux, since E[1{u < p} w / p]
 * = w). Each path is traced once and stored (face, arrival weight, sticking); with
 * w_{k+1} = c_k w_k (1 - S_k) (c_k the roulette factor or 1), d w_j / d S_k = -w_j / (1 - S_k)
 * for every later hit j, so from the last hit back
 *   Q_k = sum over j > k of gfl[f_j] w_j = gfl[f_{k+1}] w_{k+1} + Q_{k+1},
 * and gbare[f_k] += -beta Q_k / (1 - S_k). Per-face sums in double. */
void oracle_trace_neutral_rr_adjoint(void* h, const oc_box* b, const oc_source* s,
                                     const float* bare, float beta, int n, uint64_t seed,
                                     const double* gfl, double* gbare) {
  const oc_bvh* B = (const oc_bvh*)h;
  int F = B->nf;
  memset(gbare, 0, sizeof(double) * F);
#pragma omp parallel
  {
    double* acc = (double*)calloc(F > 0 ? F : 1, sizeof(double));
    int* pf = (int*)malloc(sizeof(int) * 1024);
    float* pw = (float*)malloc(sizeof(float) * 1024);
    float* pS = (float*)malloc(sizeof(float) * 1024);
#pragma omp for schedule(dynamic, 256)
    for (int p = 0; p < n; ++p) {
      oc_rng r;
      oc_rng_init(&r, seed, (uint32_t)p);
      f3 o, d, nn;
      oc_source_ray(&r, b, s, &o, &d);
      float w = 1.f, t;
      int K = 0;
      for (int bounce = 0; bounce < 1024; ++bounce) {
        int f = oc_query_bc(B, b, &o, d, &t, &nn);
        if (f < 0) break;
        if (dot3(nn, d) > 0.f) nn = scl(-1.f, nn);
        float S = bare[f] * beta;
        pf[K] = f;
        pw[K] = w;
        pS[K] = S;
        ++K;
        w = w - w * S;
        if (w <= 1.0e-6f) break;
        uint32_t u[4];
        oc_block(seed, (uint32_t)p, (uint32_t)bounce + 1u, 1u, u);
        if (w < 0.1f) {
          if (oc_u24(u[0]) < (1.f - w / 0.3f)) break;
          w = 0.3f;
        }
        f3 hit = add3(o, scl(t, d));
        d = oc_cosine_u(nn, oc_u24(u[1]), oc_u24(u[2]));
        o = add3(hit, scl(1.0e-4f, nn));
      }
      double Q = 0.0; /* sum of gfl w over the hits after k */
      for (int k = K - 1; k >= 0; --k) {
        if (k + 1 < K) Q += gfl[pf[k + 1]] * (double)pw[k + 1];
        const double q = 1.0 - (double)pS[k];
        if (k + 1 < K && q > 0.0) acc[pf[k]] += -(double)beta * Q / q;
      }
    }
#pragma omp critical
    for (int f = 0; f < F; ++f) gbare[f] += acc[f];
    free(acc);
    free(pf);
    free(pw);
    free(pS);
  }
}

/* _trace3d_ion_yield (threed.py:599-712); out = 5 x F doubles (sp, ie, ip, arr, ang). */
void oracle_trace_ion_yield(void* h, const oc_box* b, const oc_source* s, const oc_ion* P, int n,
                            uint64_t seed, double* out) {
  const oc_bvh* B = (const oc_bvh*)h;
  int F = B->nf;
  memset(out, 0, sizeof(double) * 5 * F);
  const float PI = 3.14159265f, PI_2 = 1.5707963f;
#pragma omp parallel
  {
    double* acc = (double*)calloc(F > 0 ? 5 * F : 1, sizeof(double));
#pragma omp for schedule(dynamic, 256)
    for (int p = 0; p < n; ++p) {
      oc_rng r;
      oc_rng_init(&r, seed, (uint32_t)p);
      f3 o, d, nn;
      oc_source_ray(&r, b, s, &o, &d);
      float E = P->meanE + P->sigmaE * oc_normal(&r);
      if (E < 0.f) E = 0.f;
      float w = 1.f, t;
      float sEsp = sqrtf(P->Eth_sp), sEie = sqrtf(P->Eth_ie), sEp = sqrtf(P->Eth_p);
      for (int bounce = 0; bounce < 1024; ++bounce) {
        int f = oc_query_bc(B, b, &o, d, &t, &nn);
        if (f < 0) break;
        if (dot3(nn, d) > 0.f) nn = scl(-1.f, nn);
        float cosT = -dot3(d, nn);
        if (cosT < 0.f) cosT = 0.f;
        if (cosT > 1.f) cosT = 1.f;
        float incAng = acosf(cosT);
        float sqrtE = sqrtf(E);
        float f_sp = (1.f + P->B_sp * (1.f - cosT * cosT)) * cosT;
        if (f_sp < 0.f) f_sp = 0.f;
        float f_ie = 1.f;
        if (cosT < 0.5f) {
          f_ie = 3.f - 6.f * incAng / PI;
          if (f_ie < 0.f) f_ie = 0.f;
        }
        acc[f] += P->A_sp * fmaxf(sqrtE - sEsp, 0.f) * f_sp * w;
        acc[F + f] += P->A_ie * fmaxf(sqrtE - sEie, 0.f) * f_ie * w;
        acc[2 * F + f] += P->A_p * fmaxf(sqrtE - sEp, 0.f) * f_ie * w;
        acc[3 * F + f] += w;
        acc[4 * F + f] += cosT * w;
        float stick = 1.f;
        if (incAng > P->thrmin) {
          float rr = (incAng - P->thrmin) / (P->thrmax - P->thrmin);
          if (rr < 0.f) rr = 0.f;
          if (rr > 1.f) rr = 1.f;
          stick = 1.f - rr;
        }
        if (stick >= 1.f) break;
        float Eref;
        if (incAng >= P->inflect)
          Eref = 1.f - (1.f - P->A_en) * (PI_2 - incAng) / (PI_2 - P->inflect);
        else
          Eref = P->A_en * powf(incAng / P->inflect, P->n_l);
        float newE = Eref * E;
        for (int e = 0; e < 4; ++e) {
          float cand = Eref * E + 0.1f * E * oc_normal(&r);
          if (cand >= 0.f && cand <= E) {
            newE = cand;
            break;
          }
        }
        if (newE < 0.f) newE = 0.f;
        if (newE > E) newE = E;
        if (newE <= P->minE) break;
        E = newE;
        w = w - w * stick;
        if (w <= 1.0e-4f) break;
        f3 hit = add3(o, scl(t, d));
        f3 axis = nrmz(sub3(d, scl(2.f * dot3(d, nn), nn)));
        f3 tt, bb;
        if (axis.z < -0.9999999f) {
          tt = mk(0.f, -1.f, 0.f);
          bb = mk(-1.f, 0.f, 0.f);
        } else {
          float aa = 1.f / (1.f + axis.z);
          float cc = -axis.x * axis.y * aa;
          tt = mk(1.f - axis.x * axis.x * aa, cc, -axis.x);
          bb = mk(cc, 1.f - axis.y * axis.y * aa, -axis.y);
        }
        float maxcone = PI_2 - fminf(incAng, P->minang);
        float theta = 0.f;
        for (int k = 0; k < 64; ++k) {
          float uu = sqrtf(oc_uniform(&r));
          float ss = sqrtf(fmaxf(1.f - uu, 0.f));
          theta = maxcone * ss;
          if (oc_uniform(&r) * theta * uu <= cosf(PI_2 * ss) * sinf(theta)) break;
        }
        float phi = 6.2831853f * oc_uniform(&r);
        f3 nd = nrmz(add3(scl(sinf(theta), add3(scl(cosf(phi), tt), scl(sinf(phi), bb))),
                          scl(cosf(theta), axis)));
        float dn = dot3(nd, nn);
        if (dn <= 0.f) nd = sub3(nd, scl(2.f * dn, nn));
        d = nd;
        o = add3(hit, scl(1.0e-4f, nn));
      }
    }
#pragma omp critical
    for (int f = 0; f < 5 * F; ++f) out[f] += acc[f];
    free(acc);
  }
}

/* _trace3d (threed.py:352-406); out = 2 x F doubles (flux, angacc). */
void oracle_trace_legacy(void* h, const oc_box* b, const oc_source* s, float sticking,
                         int n_reemit, int specular, float cos_thr, float eta, int n,
                         uint64_t seed, double* out) {
  const oc_bvh* B = (const oc_bvh*)h;
  int F = B->nf;
  memset(out, 0, sizeof(double) * 2 * F);
#pragma omp parallel
  {
    double* acc = (double*)calloc(F > 0 ? 2 * F : 1, sizeof(double));
#pragma omp for schedule(dynamic, 256)
    for (int p = 0; p < n; ++p) {
      oc_rng r;
      oc_rng_init(&r, seed, (uint32_t)p);
      f3 o, d, nn;
      oc_source_ray(&r, b, s, &o, &d);
      float w = 1.f, t;
      for (int bounce = 0; bounce <= n_reemit; ++bounce) {
        int f = oc_query(B, o, d, &t, &nn);
        if (f < 0) break;
        if (dot3(nn, d) > 0.f) nn = scl(-1.f, nn);
        float cosang = -dot3(d, nn);
        if (cosang < 0.f) cosang = 0.f;
        if (specular) {
          if (cosang >= cos_thr || bounce == n_reemit) {
            acc[f] += w;
            acc[F + f] += w * cosang;
            break;
          }
          float R = (cos_thr - cosang) / cos_thr;
          float dep = w * (1.f - R);
          acc[f] += dep;
          acc[F + f] += dep * cosang;
          w = w * R * eta;
          f3 hit = add3(o, scl(t, d));
          d = sub3(d, scl(2.f * dot3(d, nn), nn));
          o = add3(hit, scl(1.0e-4f, nn));
        } else {
          float u = oc_uniform(&r);
          if (u < sticking || bounce == n_reemit) {
            acc[f] += w;
            acc[F + f] += w * cosang;
            break;
          }
          f3 hit = add3(o, scl(t, d));
          d = oc_cosine(&r, nn);
          o = add3(hit, scl(1.0e-4f, nn));
        }
      }
    }
#pragma omp critical
    for (int f = 0; f < 2 * F; ++f) out[f] += acc[f];
    free(acc);
  }
}

/* _trace_ff (threed.py:908-920) with the launch of threed.py:949-955. */
void oracle_trace_ff(void* h, const oc_box* b, const float* cen, const float* gn, int nf, int n_ff,
                     uint64_t seed, int* hitface) {
  const oc_bvh* B = (const oc_bvh*)h;
  long long N = (long long)nf * n_ff;
#pragma omp parallel for schedule(dynamic, 256)
  for (long long p = 0; p < N; ++p) {
    int j = (int)(p / n_ff);
    oc_rng r;
    oc_rng_init(&r, seed, (uint32_t)p);
    f3 nn = mk(gn[3 * j], gn[3 * j + 1], gn[3 * j + 2]);
    f3 o = add3(mk(cen[3 * j], cen[3 * j + 1], cen[3 * j + 2]), scl(1.0e-3f, nn));
    f3 a = fabsf(nn.x) > 0.9f ? mk(0.f, 1.f, 0.f) : mk(1.f, 0.f, 0.f);
    f3 c = crs(a, nn);
    f3 t = scl(1.f / (sqrtf(dot3(c, c)) + 1.0e-12f), c);
    f3 bb = crs(nn, t);
    float u1 = oc_uniform(&r), u2 = oc_uniform(&r);
    float ct = sqrtf(u1), st = sqrtf(1.f - u1), ph = 6.2831853f * u2;
    f3 d = add3(add3(scl(st * cosf(ph), t), scl(st * sinf(ph), bb)), scl(ct, nn));
    float tt;
    f3 n2;
    hitface[p] = oc_query_bc(B, b, &o, d, &tt, &n2);
  }
}

/* Closest face to each point with the oracle BVH (box-distance pruning; exact in double).
 * Ties keep the first face found; the tests accept any face at the minimum distance. */
static double oc_box_d2(const oc_node* nd, const double p[3]) {
  double s = 0.0;
  for (int k = 0; k < 3; ++k) {
    double d = 0.0;
    if (p[k] < nd->lo[k]) d = nd->lo[k] - p[k];
    else if (p[k] > nd->hi[k]) d = p[k] - nd->hi[k];
    s += d * d;
  }
  return s;
}

void oracle_bvh_closest_point(void* h, const float* pts, int n, int* face_out, double* dist_out) {
  const oc_bvh* B = (const oc_bvh*)h;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    int best = -1;
    double bd = 1e300, best_d = 0.0;
    /* pass 0: the minimum squared distance; pass 1: the lowest face id within the tie band */
    for (int pass = 0; pass < 2 && B->nf > 0; ++pass) {
      const double lim = pass ? oc_tie_lim(p, bd) : 0.0;
      int stack[128], sp = 0;
      stack[sp++] = 0;
      while (sp) {
        const oc_node* nd = &B->nodes[stack[--sp]];
        if (oc_box_d2(nd, p) > (pass ? lim : bd)) continue;
        if (nd->left < 0) {
          for (int k = nd->start; k < nd->start + nd->count; ++k) {
            int f = B->order[k];
            const double* t = B->tri + 9 * f;
            double q[3];
            oc_closest_on_tri(p, t, t + 3, t + 6, q);
            double dd = (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) +
                        (p[2] - q[2]) * (p[2] - q[2]);
            if (!pass) {
              if (dd < bd) bd = dd;
            } else if (dd <= lim && (best < 0 || f < best)) {
              best = f;
              best_d = dd;
            }
          }
        } else {
          stack[sp++] = nd->left;
          stack[sp++] = nd->right;
        }
      }
    }
    face_out[i] = best;
    dist_out[i] = best >= 0 ? sqrt(best_d) : -1.0;
  }
}
