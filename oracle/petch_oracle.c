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
void oracle_closest_point_brute(const float* verts, const int* faces, int nf, const float* pts,
                                int n, int* face_out, double* dist_out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    int best = -1;
    double bd = 1e300;
    for (int f = 0; f < nf; ++f) {
      double a[3], b[3], c[3], q[3];
      oc_vert(verts, faces[3 * f], a);
      oc_vert(verts, faces[3 * f + 1], b);
      oc_vert(verts, faces[3 * f + 2], c);
      oc_closest_on_tri(p, a, b, c, q);
      double dd = (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) +
                  (p[2] - q[2]) * (p[2] - q[2]);
      if (dd < bd) {
        bd = dd;
        best = f;
      }
    }
    face_out[i] = best;
    dist_out[i] = best >= 0 ? sqrt(bd) : -1.0;
  }
}
