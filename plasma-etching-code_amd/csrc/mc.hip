// Marching cubes at iso 0 over the (nx,ny,nz) level set (replaces extract_mesh_3d /
// extract_mesh_3d_gpu, threed.py:62-100).
//
// Triangulation rule (shared with the oracle's C restatement, oracle/petch_oracle.c):
//   - node c is solid iff phi > 0;
//   - each cube face contributes the segments of the surface on that face, oriented with the
//     solid on the left seen from outside the cube; on an ambiguous face (solid corners on one
//     diagonal) the solid corners are joined iff f(s1)*f(s2) > f(g1)*f(g2) (the bilinear saddle
//     is solid), otherwise each solid corner is cut off on its own. The decision depends on the
//     face's four values only, so the two cubes sharing a face agree and the mesh is watertight;
//   - segments link into closed loops through the cube's crossing edges; loops are walked from
//     their smallest cube-edge id and triangulated by ear clipping: the first ear (in loop order)
//     whose new diagonal does not join two points of one cube face is cut (such a diagonal lies
//     in the face plane and the neighbouring cube could cut it too, leaving an edge with four
//     faces), else the first ear; triangles are wound so cross(v1-v0, v2-v0) points into the gas.
// Output order is deterministic: vertices by (node, axis) of the grid edge they sit on, faces by
// cell then loop/fan order. Vertex positions are interpolated in float64 and rounded once.
//
// Passes: count (vertices + triangles per node, per-block sums) -> one-block scan of the block
// sums -> host reads the two totals (capacity check) -> emit vertices (+ per-node vertex base)
// -> emit faces.
#include "common.h"

#include <vector>
#include "scan.h"
#include "fastdiv.h"

namespace {

constexpr int kMcBS = 256;
// A workgroup takes a tile of `rounds` x kMcBS consecutive nodes, one round of kMcBS at a time.
// Launches over very many nodes (the grids of a batched step) use tiles of up to 8 rounds, so
// they need fewer, longer workgroups; a launch over one grid keeps 256-node tiles so that it
// still fills the chip. Vertex and face ids are prefix sums over the nodes in order, so the tile
// length does not change the output.
constexpr int kMcMaxRounds = 8;

// tile length (in rounds of kMcBS) of a launch over `nodes` nodes: about 16k workgroups or more
int mc_rounds(long long nodes) {
  const long long r = nodes / ((long long)kMcBS * 16384);
  return (int)(r < 1 ? 1 : (r > kMcMaxRounds ? kMcMaxRounds : r));
}

__device__ __forceinline__ int edge_base(int e) {
  int a = e >> 2, q = e & 3, lo = q & 1, hi = q >> 1;
  if (a == 0) return (lo << 1) | (hi << 2);
  if (a == 1) return lo | (hi << 2);
  return lo | (hi << 1);
}

__device__ __forceinline__ int edge_of(int c0, int c1) {
  int d = c0 ^ c1;
  int a = (d == 1) ? 0 : ((d == 2) ? 1 : 2);
  int b = c0 & c1;
  int q;
  if (a == 0)
    q = ((b >> 1) & 1) | (((b >> 2) & 1) << 1);
  else if (a == 1)
    q = (b & 1) | (((b >> 2) & 1) << 1);
  else
    q = (b & 1) | (((b >> 1) & 1) << 1);
  return a * 4 + q;
}

// Successor table of the cell's surface loops: 4 bits per cube edge, 15 = not a crossing edge.
__device__ __forceinline__ unsigned long long cell_links(const float f[8], int& nedges) {
#pragma clang fp contract(off)
  int s = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) s |= (f[c] > 0.f ? 1 : 0) << c;
  unsigned long long nxt = ~0ull;
  nedges = 0;
  if (s == 0 || s == 255) return nxt;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int u = (a + 1) % 3, v = (a + 2) % 3;
#pragma unroll
    for (int sg = 0; sg < 2; ++sg) {
      // corner cycle, counter-clockwise seen from outside the cube
      int cyc[4];
      if (sg) {
        cyc[0] = (1 << a);
        cyc[1] = (1 << a) | (1 << u);
        cyc[2] = (1 << a) | (1 << u) | (1 << v);
        cyc[3] = (1 << a) | (1 << v);
      } else {
        cyc[0] = 0;
        cyc[1] = (1 << v);
        cyc[2] = (1 << u) | (1 << v);
        cyc[3] = (1 << u);
      }
      int sol[4];
      int ncross = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) sol[k] = (s >> cyc[k]) & 1;
#pragma unroll
      for (int k = 0; k < 4; ++k) ncross += (sol[k] != sol[(k + 1) & 3]);
      if (ncross == 0) continue;
      int E[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) E[k] = edge_of(cyc[k], cyc[(k + 1) & 3]);
      if (ncross == 2) {
        int X = 0, Y = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (sol[k] && !sol[(k + 1) & 3]) X = E[k];
          if (!sol[k] && sol[(k + 1) & 3]) Y = E[k];
        }
        nxt = (nxt & ~(15ull << (4 * X))) | ((unsigned long long)Y << (4 * X));
      } else {
        int sa = sol[0] ? 0 : 1;
        float ps = f[cyc[sa]] * f[cyc[sa + 2]];
        float pg = f[cyc[sa ^ 1]] * f[cyc[(sa ^ 1) + 2]];
        bool join = ps > pg;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (sol[k] && !sol[(k + 1) & 3]) {
            int to = join ? E[(k + 1) & 3] : E[(k + 3) & 3];
            nxt = (nxt & ~(15ull << (4 * E[k]))) | ((unsigned long long)to << (4 * E[k]));
          }
        }
      }
    }
  }
#pragma unroll
  for (int e = 0; e < 12; ++e) nedges += (((nxt >> (4 * e)) & 15) != 15);
  return nxt;
}

__device__ __forceinline__ int link_of(unsigned long long nxt, int e) {
  return (int)((nxt >> (4 * e)) & 15);
}

// Number of triangles = crossing edges - 2 * loops.
__device__ __forceinline__ int cell_tri_count(const float f[8]) {
  int ne;
  unsigned long long nxt = cell_links(f, ne);
  if (ne == 0) return 0;
  int visited = 0, loops = 0;
  for (int e0 = 0; e0 < 12; ++e0) {
    if (link_of(nxt, e0) == 15 || ((visited >> e0) & 1)) continue;
    ++loops;
    int e = e0;
    for (int g = 0; g < 12; ++g) {
      visited |= 1 << e;
      e = link_of(nxt, e);
      if (e == e0 || e == 15) break;
    }
  }
  return ne - 2 * loops;
}

struct Grid {
  int nx, ny, nz;
  long long sx, sy;  // node strides
  petch::FastDiv dy, dz;  // division by ny and nz (node indices of one grid fit 32 bits)
  int rounds;        // tile length of the launch, in rounds of kMcBS nodes
};

__host__ inline Grid make_grid(int nx, int ny, int nz, int rounds) {
  return Grid{nx, ny, nz, (long long)ny * nz, (long long)nz, petch::make_fastdiv((uint32_t)ny),
              petch::make_fastdiv((uint32_t)nz), rounds};
}

__device__ __forceinline__ void node_ijk(const Grid& g, long long n, int& i, int& j, int& k) {
  const uint32_t u = (uint32_t)n;
  const uint32_t c = petch::fastdiv(u, g.dz);   // column (i, j)
  k = (int)(u - c * (uint32_t)g.nz);
  i = (int)petch::fastdiv(c, g.dy);
  j = (int)(c - (uint32_t)i * (uint32_t)g.ny);
}

__device__ __forceinline__ long long corner_off(const Grid& g, int c) {
  return (c & 1) * g.sx + ((c >> 1) & 1) * g.sy + ((c >> 2) & 1);
}

__device__ __forceinline__ int vertex_flags(const float* __restrict__ phi, const Grid& g,
                                            long long n, int i, int j, int k) {
  bool s0 = phi[n] > 0.f;
  int fl = 0;
  if (i < g.nx - 1 && ((phi[n + g.sx] > 0.f) != s0)) fl |= 1;
  if (j < g.ny - 1 && ((phi[n + g.sy] > 0.f) != s0)) fl |= 2;
  if (k < g.nz - 1 && ((phi[n + 1] > 0.f) != s0)) fl |= 4;
  return fl;
}

__device__ __forceinline__ void load_cell(const float* __restrict__ phi, const Grid& g,
                                          long long n, float f[8]) {
#pragma unroll
  for (int c = 0; c < 8; ++c) f[c] = phi[n + corner_off(g, c)];
}

// Sign mask of the cell at node n (bit c: corner c solid), or 0 for a node that starts no cell.
__device__ __forceinline__ int cell_signs(const float* __restrict__ phi, const Grid& g,
                                          long long n, int i, int j, int k) {
  if (i >= g.nx - 1 || j >= g.ny - 1 || k >= g.nz - 1) return 0;
  int s = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) s |= (phi[n + corner_off(g, c)] > 0.f ? 1 : 0) << c;
  return s;
}

// The cells of a tile that the surface crosses (sign mask neither 0 nor 255), in node order:
// every round's crossing cells are appended to the workgroup's list `cells` (node ids, which fit
// 32 bits), and the triangle work then runs over the list with every lane on a crossing cell.
// Without the list a wavefront ran the long per-cell loop walk whenever one of its 64 nodes was
// a crossing cell, i.e. in most wavefronts, with one or two lanes active.
__device__ __forceinline__ int mc_append_cell(bool live, long long n, uint32_t* cells, int na) {
  int tot;
  const int pos = petch::block_excl_scan<kMcBS>(live ? 1 : 0, tot);
  if (live) cells[na + pos] = (uint32_t)n;
  return na + tot;
}

// The three passes over the nodes of block `blk` of one grid (blk = blockIdx.x for a single
// grid; the batched launches pass the block's index within its grid).
__device__ __forceinline__ void mc_count_block(const float* __restrict__ phi, const Grid& g,
                                               long long N, long long blk, int* __restrict__ bv,
                                               int* __restrict__ bt) {
  __shared__ uint32_t cells[kMcMaxRounds * kMcBS];
  int vc = 0, na = 0;
  for (int r = 0; r < g.rounds; ++r) {
    const long long n = (blk * g.rounds + r) * kMcBS + threadIdx.x;
    int s = 0;
    if (n < N) {
      int i, j, k;
      node_ijk(g, n, i, j, k);
      vc += __popc(vertex_flags(phi, g, n, i, j, k));
      s = cell_signs(phi, g, n, i, j, k);
    }
    na = mc_append_cell(s != 0 && s != 255, n, cells, na);
  }
  __syncthreads();
  int tc = 0;
  for (int e = threadIdx.x; e < na; e += kMcBS) {
    float f[8];
    load_cell(phi, g, (long long)cells[e], f);
    tc += cell_tri_count(f);
  }
  int sv = petch::block_sum<kMcBS>(vc);
  int st = petch::block_sum<kMcBS>(tc);
  if (threadIdx.x == 0) {
    bv[blk] = sv;
    bt[blk] = st;
  }
}

__global__ void __launch_bounds__(kMcBS) k_mc_count(const float* __restrict__ phi, Grid g,
                                                    long long N, int* __restrict__ bv,
                                                    int* __restrict__ bt) {
  mc_count_block(phi, g, N, blockIdx.x, bv, bt);
}

// The emit passes skip a block that holds no vertex (no face): its per-node vertex tables are
// never read, since a face looks up only the start nodes of its crossing edges, which hold that
// edge's vertex. nb: the grid's blocks; total: the pass's grand total (the end of the last
// block's range).
__device__ __forceinline__ bool mc_block_empty(const int* __restrict__ part, long long blk,
                                               long long nb, const int* __restrict__ total) {
  return (blk + 1 < nb ? part[blk + 1] : *total) == part[blk];
}

// One node's vertex-table entries and the vertices on its crossing edges (ids from vb on).
__device__ __forceinline__ void mc_node_verts(const float* __restrict__ phi, const Grid& g,
                                              long long n, int i, int j, int k, int fl, int vb,
                                              double dx, int* __restrict__ vbase,
                                              uint8_t* __restrict__ vflg,
                                              float* __restrict__ verts, int cap_v) {
#pragma clang fp contract(off)
  // only a node with a crossing edge is ever looked up (a face reads the start node of each of
  // its crossing edges), so the others' entries are left unwritten
  if (!fl) return;
  vbase[n] = vb;
  vflg[n] = (uint8_t)fl;
  double p0 = (double)phi[n];
  long long st[3] = {g.sx, g.sy, 1};
  int r = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!((fl >> a) & 1)) continue;
    double p1 = (double)phi[n + st[a]];
    double t = p0 / (p0 - p1);
    double c[3] = {(double)i, (double)j, (double)k};
    c[a] = c[a] + t;
    int vid = vb + r++;
    if (vid < cap_v) {
      verts[3LL * vid + 0] = (float)(c[0] * dx);
      verts[3LL * vid + 1] = (float)(c[1] * dx);
      verts[3LL * vid + 2] = (float)(c[2] * dx);
    }
  }
}

__device__ __forceinline__ void mc_verts_block(const float* __restrict__ phi, const Grid& g,
                                               long long N, long long blk, double dx,
                                               const int* __restrict__ bv,
                                               int* __restrict__ vbase,
                                               uint8_t* __restrict__ vflg,
                                               float* __restrict__ verts, int cap_v,
                                               long long nb, const int* __restrict__ total) {
#pragma clang fp contract(off)
  if (mc_block_empty(bv, blk, nb, total)) return;
  int base = bv[blk];
  for (int r = 0; r < g.rounds; ++r) {
    const long long n = (blk * g.rounds + r) * kMcBS + threadIdx.x;
    int fl = 0, i = 0, j = 0, k = 0;
    if (n < N) {
      node_ijk(g, n, i, j, k);
      fl = vertex_flags(phi, g, n, i, j, k);
    }
    int tot;
    const int vb = base + petch::block_excl_scan<kMcBS>(__popc(fl), tot);
    base += tot;
    if (n < N) mc_node_verts(phi, g, n, i, j, k, fl, vb, dx, vbase, vflg, verts, cap_v);
  }
}

__global__ void __launch_bounds__(kMcBS) k_mc_verts(const float* __restrict__ phi, Grid g,
                                                    long long N, double dx,
                                                    const int* __restrict__ bv,
                                                    int* __restrict__ vbase,
                                                    uint8_t* __restrict__ vflg,
                                                    float* __restrict__ verts, int cap_v,
                                                    long long nb, const int* __restrict__ total) {
  mc_verts_block(phi, g, N, blockIdx.x, dx, bv, vbase, vflg, verts, cap_v, nb, total);
}

// The two cube faces an edge lies on, as bits (2*axis + side) of a 6-bit mask.
__device__ __forceinline__ int edge_faces(int e) {
  int a = e >> 2, lo = edge_base(e), m = 0;
#pragma unroll
  for (int b = 0; b < 3; ++b)
    if (b != a) m |= 1 << (2 * b + ((lo >> b) & 1));
  return m;
}

__device__ __forceinline__ int edge_vertex(const Grid& g, long long n, int ce,
                                           const int* __restrict__ vbase,
                                           const uint8_t* __restrict__ vflg) {
  long long base = n + corner_off(g, edge_base(ce));
  return vbase[base] + __popc(vflg[base] & ((1 << (ce >> 2)) - 1));
}

// The triangles of the cell at node n (f: its corner values; ids from tb on).
__device__ __forceinline__ void mc_cell_faces(const Grid& g, long long n, const float f[8], int tb,
                                              const int* __restrict__ vbase,
                                              const uint8_t* __restrict__ vflg,
                                              int* __restrict__ faces, int cap_f) {
  int ne;
  unsigned long long nxt = cell_links(f, ne);
  int visited = 0, q = 0;
  for (int e0 = 0; e0 < 12; ++e0) {
    if (link_of(nxt, e0) == 15 || ((visited >> e0) & 1)) continue;
    // walk the loop once, recording its edges (4 bits each)
    unsigned long long L = 0;
    int len = 0, e = e0;
    for (int s = 0; s < 12; ++s) {
      L |= (unsigned long long)e << (4 * len);
      ++len;
      visited |= 1 << e;
      e = link_of(nxt, e);
      if (e == e0 || e == 15) break;
    }
    // ear clipping: cut the first ear whose new diagonal keeps off the cube faces
    while (len >= 3) {
      int pick = 0;
      if (len > 3) {
        for (int i = 0; i < len; ++i) {
          int ea = (int)((L >> (4 * ((i + len - 1) % len))) & 15);
          int eb = (int)((L >> (4 * ((i + 1) % len))) & 15);
          if (!(edge_faces(ea) & edge_faces(eb))) {
            pick = i;
            break;
          }
        }
      } else {
        pick = 1;
      }
      int ea = (int)((L >> (4 * ((pick + len - 1) % len))) & 15);
      int eb = (int)((L >> (4 * ((pick + 1) % len))) & 15);
      int ec = (int)((L >> (4 * pick)) & 15);
      int t = tb + q++;
      if (t < cap_f) {
        faces[3LL * t + 0] = edge_vertex(g, n, ea, vbase, vflg);
        faces[3LL * t + 1] = edge_vertex(g, n, eb, vbase, vflg);
        faces[3LL * t + 2] = edge_vertex(g, n, ec, vbase, vflg);
      }
      if (len == 3) break;
      unsigned long long lo = L & ((1ull << (4 * pick)) - 1);
      unsigned long long hi = L >> (4 * (pick + 1));
      L = lo | (hi << (4 * pick));
      --len;
    }
  }
}

__device__ __forceinline__ void mc_faces_block(const float* __restrict__ phi, const Grid& g,
                                               long long N, long long blk,
                                               const int* __restrict__ bt,
                                               const int* __restrict__ vbase,
                                               const uint8_t* __restrict__ vflg,
                                               int* __restrict__ faces, int cap_f, long long nb,
                                               const int* __restrict__ total) {
  if (mc_block_empty(bt, blk, nb, total)) return;
  __shared__ uint32_t cells[kMcMaxRounds * kMcBS];
  int na = 0;
  for (int r = 0; r < g.rounds; ++r) {
    const long long n = (blk * g.rounds + r) * kMcBS + threadIdx.x;
    int s = 0;
    if (n < N) {
      int i, j, k;
      node_ijk(g, n, i, j, k);
      s = cell_signs(phi, g, n, i, j, k);
    }
    na = mc_append_cell(s != 0 && s != 255, n, cells, na);
  }
  __syncthreads();
  // the listed cells in node order, kMcBS at a time: each one's faces start at the block's
  // offset plus the triangles of the listed cells before it (cells off the list have none)
  int base = bt[blk];
  for (int c0 = 0; c0 < na; c0 += kMcBS) {
    const int e = c0 + threadIdx.x;
    float f[8];
    int tc = 0;
    long long n = 0;
    if (e < na) {
      n = (long long)cells[e];
      load_cell(phi, g, n, f);
      tc = cell_tri_count(f);
    }
    int tot;
    const int tb = base + petch::block_excl_scan<kMcBS>(tc, tot);
    base += tot;
    if (tc) mc_cell_faces(g, n, f, tb, vbase, vflg, faces, cap_f);
  }
}

__global__ void __launch_bounds__(kMcBS) k_mc_faces(const float* __restrict__ phi, Grid g,
                                                    long long N, const int* __restrict__ bt,
                                                    const int* __restrict__ vbase,
                                                    const uint8_t* __restrict__ vflg,
                                                    int* __restrict__ faces, int cap_f,
                                                    long long nb, const int* __restrict__ total) {
  mc_faces_block(phi, g, N, blockIdx.x, bt, vbase, vflg, faces, cap_f, nb, total);
}

// The passes of many grids (the features of a batched step), one launch each: grid t owns the
// workgroups [blk0, blk0 + nb) of the node launches; one workgroup per grid scans its block
// sums. Each grid's blocks see exactly the nodes, partial sums and scans of the single call, so
// the vertices and faces are bit-identical to its output.
struct McGrid {
  const float* phi;
  Grid g;
  long long N, nb, blk0;
  double dx;
  int *bv, *bt, *totals;
  int* vbase;
  uint8_t* vflg;
  float* verts;
  int* faces;
  int cap_v, cap_f;
};

__device__ __forceinline__ int mc_grid_of(const McGrid* __restrict__ t, int n) {
  int lo = 0, hi = n - 1;
  const long long b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t[mid].blk0 <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(kMcBS) k_mc_count_many(const McGrid* __restrict__ t, int n) {
  const McGrid& m = t[mc_grid_of(t, n)];
  mc_count_block(m.phi, m.g, m.N, blockIdx.x - m.blk0, m.bv, m.bt);
}

__global__ void __launch_bounds__(1024) k_mc_scan_many(const McGrid* __restrict__ t) {
  const McGrid& m = t[blockIdx.x];
  int ca = 0, cb = 0;
  for (long long c0 = 0; c0 < m.nb; c0 += 1024) {
    const long long i = c0 + threadIdx.x;
    const int va = (i < m.nb) ? m.bv[i] : 0;
    const int vb = (i < m.nb) ? m.bt[i] : 0;
    int ta, tb;
    const int ea = petch::block_excl_scan<1024>(va, ta);
    const int eb = petch::block_excl_scan<1024>(vb, tb);
    if (i < m.nb) {
      m.bv[i] = ca + ea;
      m.bt[i] = cb + eb;
    }
    ca += ta;
    cb += tb;
  }
  if (threadIdx.x == 0) {
    m.totals[0] = ca;
    m.totals[1] = cb;
  }
}

__global__ void __launch_bounds__(kMcBS) k_mc_verts_many(const McGrid* __restrict__ t, int n) {
  const McGrid& m = t[mc_grid_of(t, n)];
  mc_verts_block(m.phi, m.g, m.N, blockIdx.x - m.blk0, m.dx, m.bv, m.vbase, m.vflg, m.verts,
                 m.cap_v, m.nb, m.totals);
}

__global__ void __launch_bounds__(kMcBS) k_mc_faces_many(const McGrid* __restrict__ t, int n) {
  const McGrid& m = t[mc_grid_of(t, n)];
  mc_faces_block(m.phi, m.g, m.N, blockIdx.x - m.blk0, m.bt, m.vbase, m.vflg, m.faces, m.cap_f,
                 m.nb, m.totals + 1);
}

}  // namespace

extern "C" int petch_marching_cubes(petch_ctx* ctx, const float* phi, int nx, int ny, int nz,
                                    double dx, float* verts, int* faces, int cap_v, int cap_f,
                                    int* nv_host, int* nf_host) {
  PETCH_ARG(ctx, ctx && phi && nv_host && nf_host, "marching_cubes: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "marching_cubes: bad shape %d,%d,%d", nx, ny, nz);
  PETCH_ARG(ctx, cap_v >= 0 && cap_f >= 0, "marching_cubes: negative capacity");
  *nv_host = 0;
  *nf_host = 0;
  if (nx < 2 || ny < 2 || nz < 2) return PETCH_OK;  // no cells, no surface
  long long N = (long long)nx * ny * nz;
  PETCH_ARG(ctx, N <= 0xffffffffLL, "marching_cubes: %lld nodes exceed 32-bit node indices", N);
  const int rounds = mc_rounds(N);
  const long long tile = (long long)rounds * kMcBS;
  Grid g = make_grid(nx, ny, nz, rounds);
  long long nb = (N + tile - 1) / tile;
  int* blk = (int*)petch_scratch(ctx, SCR_MC_BLOCK, (size_t)(2 * nb + 2) * sizeof(int));
  if (!blk) return PETCH_E_HIP;
  int* bv = blk;
  int* bt = blk + nb;
  int* totals = blk + 2 * nb;
  hipLaunchKernelGGL(k_mc_count, dim3((unsigned)nb), dim3(kMcBS), 0, ctx->stream, phi, g, N, bv,
                     bt);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(petch::k_scan_partials2, dim3(1), dim3(1024), 0, ctx->stream, bv, bt,
                     (int)nb, totals);
  PETCH_LAUNCH_CHECK(ctx);
  int* h = (int*)petch_pinned(ctx, 2 * sizeof(int));
  if (!h) return PETCH_E_HIP;
  PETCH_HIP(ctx, hipMemcpyAsync(h, totals, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  PETCH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  int nv = h[0], nf = h[1];
  *nv_host = nv;
  *nf_host = nf;
  if (nv > cap_v || nf > cap_f || (nv > 0 && !verts) || (nf > 0 && !faces))
    return petch_set_error(ctx, PETCH_E_CAPACITY,
                           "marching_cubes: need %d vertices / %d faces, capacity %d / %d", nv,
                           nf, cap_v, cap_f);
  if (nf == 0 && nv == 0) return PETCH_OK;
  unsigned char* cell = (unsigned char*)petch_scratch(ctx, SCR_MC_CELL, (size_t)N * 5);
  if (!cell) return PETCH_E_HIP;
  int* vbase = (int*)cell;
  uint8_t* vflg = cell + 4 * N;
  hipLaunchKernelGGL(k_mc_verts, dim3((unsigned)nb), dim3(kMcBS), 0, ctx->stream, phi, g, N, dx,
                     bv, vbase, vflg, verts, cap_v, nb, totals);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_mc_faces, dim3((unsigned)nb), dim3(kMcBS), 0, ctx->stream, phi, g, N, bt,
                     vbase, vflg, faces, cap_f, nb, totals + 1);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_marching_cubes_batch(petch_ctx* ctx, int n, const petch_mc_item* items) {
  PETCH_ARG(ctx, ctx && n >= 0 && (n == 0 || items), "marching_cubes_batch: null pointer or n < 0");
  if (n == 0) return PETCH_OK;
  // block partial sums of every grid side by side, then the totals of every grid
  std::vector<McGrid> t;
  std::vector<int> which;
  t.reserve((size_t)n);
  long long parts = 0, blocks = 0, nodes = 0;
  for (int f = 0; f < n; ++f) {
    const petch_mc_item& a = items[f];
    PETCH_ARG(ctx, a.phi && a.nv_out && a.nf_out, "marching_cubes_batch: item %d has a null pointer", f);
    PETCH_ARG(ctx, a.nx > 0 && a.ny > 0 && a.nz > 0, "marching_cubes_batch: item %d has a bad shape", f);
    PETCH_ARG(ctx, a.cap_v >= 0 && a.cap_f >= 0, "marching_cubes_batch: item %d has a negative capacity", f);
    *a.nv_out = 0;
    *a.nf_out = 0;
    if (a.nx < 2 || a.ny < 2 || a.nz < 2) continue;       // no cells, no surface
    McGrid m{};
    m.phi = a.phi;
    m.g = make_grid(a.nx, a.ny, a.nz, 1);
    m.N = (long long)a.nx * a.ny * a.nz;
    PETCH_ARG(ctx, m.N <= 0xffffffffLL,
              "marching_cubes_batch: item %d: %lld nodes exceed 32-bit node indices", f, m.N);
    m.dx = a.dx;
    m.verts = a.verts;
    m.faces = a.faces;
    m.cap_v = a.cap_v;
    m.cap_f = a.cap_f;
    t.push_back(m);
    which.push_back(f);
    nodes += m.N;
  }
  const int m_ = (int)t.size();
  if (m_ == 0) return PETCH_OK;
  // one tile length for the launch, from all its nodes; each grid's tiles and block-sum slots
  const int rounds = mc_rounds(nodes);
  const long long tile = (long long)rounds * kMcBS;
  for (McGrid& m : t) {
    m.g.rounds = rounds;
    m.nb = (m.N + tile - 1) / tile;
    m.blk0 = blocks;
    m.bv = (int*)(intptr_t)parts;                           // offsets until the buffer exists
    parts += 2 * m.nb;
    blocks += m.nb;
  }
  PETCH_ARG(ctx, blocks <= 0x7fffffffLL, "marching_cubes_batch: too many nodes for one launch");
  int* blk = (int*)petch_scratch(ctx, SCR_MC_BLOCK, (size_t)(parts + 2 * m_) * sizeof(int));
  int* h = (int*)petch_pinned(ctx, (size_t)(2 * m_) * sizeof(int));
  if (!blk || !h) return PETCH_E_HIP;
  for (int k = 0; k < m_; ++k) {
    McGrid& m = t[k];
    m.bv = blk + (intptr_t)m.bv;
    m.bt = m.bv + m.nb;
    m.totals = blk + parts + 2 * k;
  }
  const McGrid* d = (const McGrid*)petch_upload_table(ctx, t.data(), t.size() * sizeof(McGrid));
  if (!d) return PETCH_E_HIP;
  hipLaunchKernelGGL(k_mc_count_many, dim3((unsigned)blocks), dim3(kMcBS), 0, ctx->stream, d, m_);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_mc_scan_many, dim3((unsigned)m_), dim3(1024), 0, ctx->stream, d);
  PETCH_LAUNCH_CHECK(ctx);
  // the one host read of the batch: every grid's vertex and face count
  PETCH_HIP(ctx, hipMemcpyAsync(h, blk + parts, (size_t)(2 * m_) * sizeof(int),
                                hipMemcpyDeviceToHost, ctx->stream));
  PETCH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  int short_of = -1;
  for (int k = 0; k < m_; ++k) {
    const petch_mc_item& a = items[which[k]];
    *a.nv_out = h[2 * k];
    *a.nf_out = h[2 * k + 1];
    if (h[2 * k] > a.cap_v || h[2 * k + 1] > a.cap_f || (h[2 * k] > 0 && !a.verts) ||
        (h[2 * k + 1] > 0 && !a.faces))
      short_of = k;
  }
  if (short_of >= 0)
    return petch_set_error(ctx, PETCH_E_CAPACITY,
                           "marching_cubes_batch: item %d needs %d vertices / %d faces, capacity "
                           "%d / %d (every item's counts are in its nv_out / nf_out)",
                           which[short_of], h[2 * short_of], h[2 * short_of + 1],
                           items[which[short_of]].cap_v, items[which[short_of]].cap_f);
  // the emit passes over the grids with a surface; each gets its own per-node vertex tables
  std::vector<McGrid> e;
  e.reserve(t.size());
  long long cells = 0;
  blocks = 0;
  for (int k = 0; k < m_; ++k) {
    if (h[2 * k] == 0 && h[2 * k + 1] == 0) continue;
    McGrid m = t[k];
    m.blk0 = blocks;
    blocks += m.nb;
    m.vbase = (int*)(intptr_t)cells;
    cells += 5 * ((m.N + 63) / 64 * 64);
    e.push_back(m);
  }
  if (e.empty()) return PETCH_OK;
  unsigned char* cell = (unsigned char*)petch_scratch(ctx, SCR_MC_CELL, (size_t)cells);
  if (!cell) return PETCH_E_HIP;
  for (McGrid& m : e) {
    unsigned char* base = cell + (intptr_t)m.vbase;
    m.vbase = (int*)base;
    m.vflg = base + 4 * ((m.N + 63) / 64 * 64);
  }
  d = (const McGrid*)petch_upload_table(ctx, e.data(), e.size() * sizeof(McGrid));
  if (!d) return PETCH_E_HIP;
  const int ne = (int)e.size();
  hipLaunchKernelGGL(k_mc_verts_many, dim3((unsigned)blocks), dim3(kMcBS), 0, ctx->stream, d, ne);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_mc_faces_many, dim3((unsigned)blocks), dim3(kMcBS), 0, ctx->stream, d, ne);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}
