// Ray-traced flux kernels. Each ray owns its Philox stream (seed, ray) and generates its source
// ray in registers (no origin/direction buffers). The RR-neutral and faithful-ion kernels run a
// persistent grid over a ray queue and advance every lane's BVH query one wide node per loop
// iteration; hits are handled, and lanes whose rays ended refilled, in batches, so neither
// divergent query lengths nor divergent bounce and wrap counts idle the wavefront for long.
// Deposits go to per-face float accumulators with hardware float atomics.
//
//   k_trace_neutral_rr   _trace3d_cov_rr    threed.py:549-596 (RR-weighted neutral, <= 1024 bounces)
//   k_trace_ion_yield    _trace3d_ion_yield threed.py:599-712 (faithful ion, 5 deposits per hit)
//   k_trace_legacy       _trace3d           threed.py:352-406 (sticking neutral / specular ion)
//   k_trace_ff           _trace_ff          threed.py:908-920 (first hit of face-emitted rays)
//   k_gen_source         _gen_source        threed.py:441-458 (same streams, for checking)
//   k_trace_*_batch      the same two traces for many features from one ray queue
// Periodic lateral boundaries follow _apply_bc (threed.py:289-349); the hit found inside the wrap
// loop is the hit of the caller's query (the reference repeats the same query), so it is reused.
#include "common.h"
#include "bvh.h"
#include "rng.h"

#include <vector>

using petch::BvhView;
using petch::Hit;
using petch::Rng;
using petch::Vec3;
using petch::v3;

petch::BvhView petch_bvh_view(const petch_bvh* b);

namespace {

// Per-context 64-bit hit counters (slot 0 neutral, 1 ion), read by petch_trace_hit_counts;
// slots 2-3 hold the ray-queue cursor of the persistent trace launches.
unsigned long long* petch_hit_counter(petch_ctx* ctx, int which) {
  bool fresh = ctx->scr_cap[SCR_TRACE] == 0;
  unsigned long long* c =
      (unsigned long long*)petch_scratch(ctx, SCR_TRACE, 4 * sizeof(unsigned long long));
  if (c && fresh && hipMemsetAsync(c, 0, 4 * sizeof(unsigned long long), ctx->stream) != hipSuccess)
    return nullptr;
  return c ? c + which : nullptr;
}

constexpr int kBS = 256;
constexpr float kPi = 3.14159265f;
constexpr float kPi2 = 1.5707963f;
constexpr float kTwoPi = 6.2831853f;

struct Box {
  float lx, ly, lz, zsrc;
  int periodic;
};

enum SourceKind { SRC_NEUTRAL = 0, SRC_ION_GAUSS = 1, SRC_ION_POWCOS = 2 };

struct Source {
  int kind;
  float sigma;     // Gaussian tilt (rad)
  float exponent;  // power-cosine exponent
};

// Source ray on the plane z = zsrc, uniform in [0,Lx) x [0,Ly): ion = Gaussian tilt
// (threed.py:450-453) or power-cosine cone about -z; neutral = cosine hemisphere about -z
// (threed.py:454-457).
__device__ __forceinline__ void gen_source(Rng& r, const Box& b, const Source& s, Vec3& o,
                                           Vec3& d) {
  float ox = r.uniform() * b.lx;
  float oy = r.uniform() * b.ly;
  o = v3(ox, oy, b.zsrc);
  if (s.kind == SRC_ION_GAUSS) {
    float ax = r.normal() * s.sigma;
    float ay = r.normal() * s.sigma;
    d = petch::normalize(v3(sinf(ax), sinf(ay), -cosf(ax) * cosf(ay)));
  } else if (s.kind == SRC_ION_POWCOS) {
    // cos(theta) = u^(1/(n+1)); 1 - cos from expm1 keeps sin(theta) accurate near the axis
    float L = logf(fmaxf(r.uniform(), 1.0e-30f)) / (s.exponent + 1.f);
    float ct = expf(L);
    float st = sqrtf(fmaxf(-expm1f(L) * (1.f + ct), 0.f));
    float ph = kTwoPi * r.uniform();
    d = petch::normalize(v3(st * cosf(ph), st * sinf(ph), -ct));
  } else {
    float r1 = r.uniform(), r2 = r.uniform();
    float ct = sqrtf(r1), st = sqrtf(1.f - r1), ph = kTwoPi * r2;
    d = petch::normalize(v3(st * cosf(ph), st * sinf(ph), -ct));
  }
}

// Tangent frame (t, b) of the cosine lobe about the unit normal n (threed.py:395-405).
__device__ __forceinline__ void cosine_frame(Vec3 n, Vec3& t, Vec3& b) {
  Vec3 a = (fabsf(n.x) > 0.9f) ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
  t = petch::normalize(petch::cross(a, n));
  b = petch::cross(n, t);
}

// Cosine-weighted direction about the unit normal n (threed.py:395-405).
__device__ __forceinline__ Vec3 cosine_about(Rng& r, Vec3 n) {
  Vec3 t, bb;
  cosine_frame(n, t, bb);
  float r1 = r.uniform(), r2 = r.uniform();
  float ct = sqrtf(r1), st = sqrtf(1.f - r1), ph = kTwoPi * r2;
  return (st * cosf(ph)) * t + (st * sinf(ph)) * bb + ct * n;
}

// Direction request of a lane that is about to start a query: the cosine lobe about the unit
// normal n (a re-emission, kLobe), the neutral source's lobe about -z, normalised as gen_source
// does (kSourceLobe), or a direction already drawn, held in n (kDirect). r1, r2 are the lobe's
// two uniforms. The trace loop draws every lane's direction at one place, so the frame, square
// roots, sine and cosine run once per refill for all lanes, not once per code path.
enum EmitMode { kLobe = 0, kSourceLobe = 1, kDirect = 2 };
struct Emit {
  Vec3 n;
  float r1, r2;
  int mode;
};

__device__ __forceinline__ Vec3 emit_dir(const Emit& e) {
  if (e.mode == kDirect) return e.n;
  const float ct = sqrtf(e.r1), st = sqrtf(1.f - e.r1), ph = kTwoPi * e.r2;
  const float c = cosf(ph), s = sinf(ph);
  if (e.mode == kSourceLobe) return petch::normalize(v3(st * c, st * s, -ct));
  Vec3 t, b;
  cosine_frame(e.n, t, b);
  return (st * c) * t + (st * s) * b + ct * e.n;
}

// One wrap of _apply_bc (threed.py:289-349) for a ray that missed the surface: if it leaves
// through a lateral face of the box before the open top it re-enters at the opposite face with
// the same direction (origin updated, returns true); otherwise it escapes (returns false).
// The exit face is chosen by comparing the distances to the three faces computed with the
// direction's reciprocals inv (a ray's constants already hold them); only the chosen face's
// distance is then divided exactly, as _apply_bc computes it, for the re-entry point. The
// choice can differ from comparing three exact quotients only where two faces are reached
// within an ulp of each other (a corner exit).
__device__ __forceinline__ bool wrap_ray(const Box& b, Vec3& o, Vec3 d, Vec3 inv) {
  float tx = 1.0e30f, ty = 1.0e30f, tz = 1.0e30f;
  const float ex = d.x > 0.f ? b.lx : 0.f, ey = d.y > 0.f ? b.ly : 0.f;
  if (fabsf(d.x) > 1.0e-9f) tx = (ex - o.x) * inv.x;
  if (fabsf(d.y) > 1.0e-9f) ty = (ey - o.y) * inv.y;
  if (d.z > 1.0e-9f) tz = (b.lz - o.z) * inv.z;
  if (tx < tz && tx < ty && tx > 1.0e-7f) {
    tx = (ex - o.x) / d.x;
    float hy = o.y + tx * d.y, hz = o.z + tx * d.z;
    o = v3(d.x < 0.f ? b.lx - 1.0e-4f : 1.0e-4f, hy, hz);
    return true;
  }
  if (ty < tz && ty < 1.0e29f && ty > 1.0e-7f) {
    ty = (ey - o.y) / d.y;
    float hx = o.x + ty * d.x, hz = o.z + ty * d.z;
    o = v3(hx, d.y < 0.f ? b.ly - 1.0e-4f : 1.0e-4f, hz);
    return true;
  }
  return false;
}

__device__ __forceinline__ bool wrap_ray(const Box& b, Vec3& o, Vec3 d) {
  return wrap_ray(b, o, d, v3(1.f / d.x, 1.f / d.y, 1.f / d.z));
}

constexpr int kMaxWraps = 256;

// A ray below all triangles and not rising (or above all and not falling) can meet no triangle
// again: a wrap moves it only in x or y. _apply_bc would wrap it until it escapes through the
// top or runs out of wraps, with every query a miss, so ending it at once gives the same result.
// On a trench whose mesh stops one cell short of the periodic faces (the grid spans
// [0, (n-1) dx], the period n dx), the rays that slip through that strip fall below the floor and
// would spend the 256 wraps there: over half of all queries of a neutral trace.
__device__ __forceinline__ bool never_returns(const BvhView& bv, Vec3 o, Vec3 d) {
  if (!bv.root) return false;
  return (o.z < bv.root[2] && d.z <= 0.f) || (o.z > bv.root[5] && d.z >= 0.f);
}

// _apply_bc + mesh_query_ray: with periodic lateral faces a ray that misses the surface and
// leaves through x or y before the open top re-enters at the opposite face (<= 256 wraps, then
// one last query whose result stands).
// A query whose ray cannot meet the box of all triangles is skipped (it would miss).
__device__ __forceinline__ Hit trace_bc(const BvhView& bv, const Box& b, Vec3& o, Vec3 d) {
  if (!b.periodic) return petch::bvh_closest_hit(bv, o, d, 1.0e6f);
  for (int w = 0; w < kMaxWraps; ++w) {
    const petch::RayPre r = petch::ray_pre(o, d);
    if (petch::root_may_hit(bv, r)) {
      petch::ScratchStack stk;
      Hit h = petch::bvh_closest_hit(bv, r, 1.0e6f, stk);
      if (h.face >= 0) return h;
    }
    if (never_returns(bv, o, d) || !wrap_ray(b, o, d)) return Hit{-1, -1, 1.0e6f};
  }
  return petch::bvh_closest_hit(bv, o, d, 1.0e6f);
}

__device__ __forceinline__ Vec3 hit_normal(const BvhView& bv, int slot) {
  float4 q0 = bv.tris[3 * slot], q1 = bv.tris[3 * slot + 1], q2 = bv.tris[3 * slot + 2];
  Vec3 a = v3(q0.x, q0.y, q0.z);
  return petch::normalize(petch::cross(v3(q1.x, q1.y, q1.z) - a, v3(q2.x, q2.y, q2.z) - a));
}

// Per-face accumulation: a float atomic add, or with `fixed` (the deterministic mode) a 64-bit
// integer add of v in units of 2^-32 into an int64 accumulator at the same address. Integer sums
// do not depend on the order of the adds, so the result is the same whatever the schedule; it is
// the exact sum of the deposits rounded to 2^-32 (|sum| < 2^31 per face).
constexpr double kFixScale = 4294967296.0;
__device__ __forceinline__ void deposit(float* __restrict__ acc, int face, float v, int fixed) {
  if (fixed)
    atomicAdd((unsigned long long*)acc + face,
              (unsigned long long)__double2ll_rn((double)v * kFixScale));
  else
    unsafeAtomicAdd(acc + face, v);
}
__device__ __forceinline__ void deposit(float* __restrict__ acc, int face, float v) {
  unsafeAtomicAdd(acc + face, v);
}

// Adds the lane's hit count to a 64-bit device counter with one atomic per wavefront.
__device__ __forceinline__ void count_hits(unsigned long long* __restrict__ ctr, unsigned hits) {
  if (!ctr) return;
  unsigned long long v = hits;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(ctr, v);
}

// Ray queue shared by a persistent grid: each wavefront refills its finished lanes with new ray
// ids, so lanes whose rays bounce many times deep in a feature do not hold the whole wavefront
// hostage to its longest ray. A wavefront takes ids from the global counter in chunks of
// kRayChunk and hands them out from a private pool (its `pool`, `pool_end` are uniform across
// the wavefront): the counter is one device-scope atomic location that every XCD hits, and
// one atomic per refill (about 40 rays) was what bounded the trace kernels. Measured on the
// AR-10 trench (8M neutral rays): 6.05 ms per launch with an atomic per refill, 4.29 / 4.35 /
// 4.68 ms with chunks of 128 / 256 / 512 (larger chunks leave the last ones unevenly spread).
// `exhausted` is uniform across the wavefront: set when a chunk starts at or beyond the end,
// after the pool's last ids were handed out.
struct RayQueue {
  unsigned* next;
  int n;
};

constexpr unsigned kRayChunk = 128;  // >= 64, so one chunk always covers a wavefront's refill

__device__ __forceinline__ bool claim_ray(const RayQueue& q, bool active, bool& exhausted,
                                          unsigned& pool, unsigned& pool_end, unsigned& ray) {
  if (exhausted) return false;
  const int lane = threadIdx.x & 63;
  const unsigned long long idle = __ballot(!active);
  if (!idle) return false;
  const unsigned cnt = (unsigned)__popcll(idle);
  const unsigned left = pool_end - pool;
  const unsigned long long below = (lane == 0) ? 0ull : (idle & ((1ull << lane) - 1ull));
  const unsigned rank = (unsigned)__popcll(below);
  unsigned id;
  if (left >= cnt) {
    id = pool + rank;
    pool += cnt;
  } else {
    // the pool's remaining ids go to the first lanes, a fresh chunk serves the rest
    const int leader = __ffsll((unsigned long long)idle) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(q.next, kRayChunk);
    base = __shfl(base, leader, 64);
    id = rank < left ? pool + rank : base + (rank - left);
    pool = base + (cnt - left);
    pool_end = base + kRayChunk;
    if (base >= (unsigned)q.n) exhausted = true;
  }
  if (active) return false;
  ray = id;
  return ray < (unsigned)q.n;
}

// Persistent trace loop shared by the RR-neutral and faithful-ion kernels. Each lane is idle,
// traversing (one wide BVH node per loop iteration) or holding a surface hit. A query that
// ends in a miss is handled on the spot: a periodic wrap (_apply_bc) moves the origin and
// restarts the query, otherwise the ray escapes and the lane goes idle. Hits are handled in
// batches: only when at most kRefillAt lanes of the wavefront are still traversing does it run
// the policy's hit code and refill idle lanes from the ray queue, so the (short, very uneven)
// queries keep most lanes busy. The policy P supplies the surface and the per-ray payload:
// view() / box() give the BVH and tracer box the ray lives in, start() keys the ray's Philox
// stream and draws its source ray's origin, hit() deposits and either ends the ray (false) or
// moves the origin to the re-emission point (true); both leave the direction as an Emit
// request that the loop turns into a direction at one place for all lanes. Every draw of a ray
// comes from that ray's own Philox stream in a fixed order, so the result does not depend on
// the schedule.
constexpr int kRefillAt = 24;
constexpr int kLdsStack = 8;  // traversal stack entries per lane held in LDS (16 KB/workgroup)
// kMiss: the ray left the domain and the policy replays paths (P::kReplay), so it gets a say
// (P::miss) before the lane goes idle.
enum LaneState { kIdle = 0, kTrav = 1, kHit = 2, kMiss = 3 };

template <class P>
__device__ __forceinline__ void persistent_trace(const P& pol, const RayQueue& q,
                                                 unsigned long long* __restrict__ hits) {
  __shared__ int lds_n[kLdsStack * kBS];
  __shared__ float lds_t[kLdsStack * kBS];
  petch::TravStack<kBS, kLdsStack> stk;
  stk.lds_n = lds_n + threadIdx.x;
  stk.lds_t = lds_t + threadIdx.x;
  petch::Trav T;
  typename P::Ray R;
  Vec3 o = v3(0.f, 0.f, 0.f), d = v3(0.f, 0.f, -1.f);
  petch::RayPre rp;  // per-direction constants of the current ray; a wrap only moves its origin
  int wraps = 0;
  unsigned nh = 0;
  int st = kIdle;
  bool exhausted = false;
  unsigned pool = 0, pool_end = 0;  // this wavefront's unclaimed queue ids [pool, pool_end)
  while (true) {
    if (__popcll(__ballot(st == kTrav)) <= kRefillAt) {
      Emit e;
      bool emit = false;
      if (st == kHit) {
        ++nh;
        emit = pol.hit(R, T.h, o, d, e);
        st = kIdle;
      }
      if constexpr (P::kReplay) {
        if (st == kMiss) {
          emit = pol.miss(R, o, e);
          st = kIdle;
        }
      }
      unsigned ray;
      if (claim_ray(q, emit || st == kTrav, exhausted, pool, pool_end, ray)) {
        pol.start(R, ray, o, e);
        emit = true;
      }
      if (emit) {
        d = emit_dir(e);
        rp = petch::ray_pre(o, d);
        wraps = 0;
        petch::trav_begin(pol.view(R), T, rp, 1.0e6f);
        st = kTrav;
      }
      if (!__ballot(st == kTrav)) break;
    }
    if (st == kTrav && petch::trav_step(pol.view(R), rp, T, stk)) {
      const Box& b = pol.box(R);
      if (T.h.face >= 0) {
        st = kHit;
      } else if (b.periodic && wraps < kMaxWraps && !never_returns(pol.view(R), o, d) &&
                 wrap_ray(b, o, d, rp.inv)) {
        petch::ray_move(rp, o);
        ++wraps;
        petch::trav_begin(pol.view(R), T, rp, 1.0e6f);
      } else {
        st = P::kReplay ? kMiss : kIdle;
      }
    }
  }
  count_hits(hits, nh);
}

// _trace3d_cov_rr (threed.py:549-596): deposit w at every hit, w *= 1 - bare*beta, Russian
// roulette below 0.1 (survivors carry 0.3), cosine re-emission, <= 1024 bounces. One launch
// can carry two species over the same surface (the F and O traces of one fixed-point
// iteration, threed.py:876-880): queue ids [0, n0) are species 0 with stream (seed[0], id),
// the rest species 1 with stream (seed[1], id - n0), exactly the streams of two launches.
//
// Stream of a neutral: its source ray takes the words of block 0 (lane 0) in order (origin x,
// origin y, then the lobe's two uniforms); its k-th bounce (k = 1, 2, ...) takes the words of
// block (k, ray, 1): word 0 for the roulette, words 1 and 2 for the re-emission lobe. A fixed
// block per event means the trace keeps no buffered RNG state and evaluates Philox at two
// places of its refill code (hits, new rays) instead of at every draw.
//
// One Russian-roulette bounce (threed.py:565-596) of a ray with weight w that hit face h:
// deposits w, applies the sticking S, plays the roulette; returns false when the ray ends, else
// moves o to the re-emission point and leaves the cosine lobe about the oriented normal in e.
// The weight step of one bounce, shared by the trace and its adjoint so that both take the same
// decisions: from weight w and sticking S at the ray's bounce `bounce`, w2 = w (1 - S), or 0.3
// after a survived roulette (*reset); false when the ray ends (w2 <= 1e-6, roulette, or the
// 1024-bounce cap), else ++bounce. u receives the bounce's Philox words.
__device__ __forceinline__ bool rr_weight(float w, float S, uint64_t seed, unsigned ray,
                                          int& bounce, float& w2, bool& reset, uint32_t u[4]) {
  w2 = w - w * S;
  reset = false;
  if (w2 <= 1.0e-6f) return false;
  petch::philox_block(seed, ray, (uint32_t)bounce + 1u, 1u, u);
  if (w2 < 0.1f) {
    if (petch::u24(u[0]) < (1.f - w2 / 0.3f)) return false;
    w2 = 0.3f;
    reset = true;
  }
  if (++bounce >= 1024) return false;
  return true;
}

// Re-emission of a continuing ray: o moves to 1e-4 off the hit point along the oriented normal
// nn, and e holds the cosine lobe about nn from the bounce's words 1 and 2.
__device__ __forceinline__ void rr_emit(const Hit& h, Vec3 nn, const uint32_t u[4], Vec3& o,
                                        Vec3 d, Emit& e) {
  const Vec3 p = o + h.t * d;
  e.n = nn;
  e.r1 = petch::u24(u[1]);
  e.r2 = petch::u24(u[2]);
  e.mode = kLobe;
  o = p + 1.0e-4f * nn;
}

__device__ __forceinline__ bool rr_bounce(float& w, int& bounce, uint64_t seed, unsigned ray,
                                          const BvhView& bv, const Hit& h, float* __restrict__ fl,
                                          int fixed, float S, Vec3& o, Vec3 d, Emit& e) {
  Vec3 nn = hit_normal(bv, h.slot);
  if (petch::dot(nn, d) > 0.f) nn = -nn;
  deposit(fl, h.face, w, fixed);
  float w2;
  bool reset;
  uint32_t u[4];
  if (!rr_weight(w, S, seed, ray, bounce, w2, reset, u)) return false;
  w = w2;
  rr_emit(h, nn, u, o, d, e);
  return true;
}

// Source ray of a neutral's stream: the cosine source (threed.py:454-457) draws its origin here
// and leaves its lobe to the loop; another source kind is drawn whole by gen_source.
__device__ __forceinline__ void neutral_source(uint64_t seed, unsigned ray, const Box& b,
                                               const Source& src, Vec3& o, Emit& e) {
  if (src.kind == SRC_NEUTRAL) {
    uint32_t u[4];
    petch::philox_block(seed, ray, 0u, 0u, u);
    o = v3(petch::u24(u[0]) * b.lx, petch::u24(u[1]) * b.ly, b.zsrc);
    e.r1 = petch::u24(u[2]);
    e.r2 = petch::u24(u[3]);
    e.mode = kSourceLobe;
  } else {
    Rng r;
    r.init(seed, ray);
    gen_source(r, b, src, o, e.n);
    e.mode = kDirect;
  }
}

struct NeutralRR {
  static constexpr bool kReplay = false;
  BvhView bv;
  Box b;
  Source src;
  const float* __restrict__ bare;
  float beta[2];
  float* fl[2];
  uint64_t seed[2];
  unsigned n0;
  int fixed;  // fl are int64 fixed-point accumulators (deterministic mode)
  struct Ray {
    float w;
    int bounce;
    int sp;
    unsigned ray;  // id within the species' stream
  };
  __device__ __forceinline__ const BvhView& view(const Ray&) const { return bv; }
  __device__ __forceinline__ const Box& box(const Ray&) const { return b; }
  __device__ __forceinline__ void start(Ray& R, unsigned ray, Vec3& o, Emit& e) const {
    R.sp = ray >= n0;
    R.ray = R.sp ? ray - n0 : ray;
    neutral_source(R.sp ? seed[1] : seed[0], R.ray, b, src, o, e);
    R.w = 1.f;
    R.bounce = 0;
  }
  __device__ __forceinline__ bool hit(Ray& R, const Hit& h, Vec3& o, Vec3 d, Emit& e) const {
    const float S = bare[h.face] * (R.sp ? beta[1] : beta[0]);
    return rr_bounce(R.w, R.bounce, R.sp ? seed[1] : seed[0], R.ray, bv, h,
                     R.sp ? fl[1] : fl[0], fixed, S, o, d, e);
  }
};

// Etch-product redeposition (mc_redep_3d, threed.py:1582-1605): product leaves face f with
// probability w_f / sum w, w_f = max(rate_f area_f, 0), from 1e-3 off its centroid along its
// into-gas normal, in the cosine lobe about that normal, then follows the RR neutral walk with
// sticking s_redep on every face (bare = 1) in an open box (no lateral wrap). cum holds the
// inclusive float64 prefix sums of w; a ray draws its face by bisection for the first prefix
// above u * total, u a 53-bit uniform from words 0 and 1 of its block 0, and its lobe from
// words 2 and 3. With nothing to emit (total <= 0) every ray carries weight 0 and deposits
// nothing; the result is then zero.
struct Redep {
  static constexpr bool kReplay = false;
  BvhView bv;
  Box b;
  const float* __restrict__ cen;
  const float* __restrict__ gn;
  const double* __restrict__ cum;
  int nf;
  float S;
  float* fl;
  uint64_t seed;
  int fixed;
  struct Ray {
    float w;
    int bounce;
    unsigned ray;
  };
  __device__ __forceinline__ const BvhView& view(const Ray&) const { return bv; }
  __device__ __forceinline__ const Box& box(const Ray&) const { return b; }
  __device__ __forceinline__ void start(Ray& R, unsigned ray, Vec3& o, Emit& e) const {
    uint32_t u[4];
    petch::philox_block(seed, ray, 0u, 0u, u);
    const double tot = cum[nf - 1];
    const double x = ((double)(u[0] >> 5) * 67108864.0 + (double)(u[1] >> 6)) *
                     (1.0 / 9007199254740992.0);
    const double target = x * tot;
    int lo = 0, hi = nf - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cum[mid] > target)
        hi = mid;
      else
        lo = mid + 1;
    }
    const Vec3 n = v3(gn[3 * lo], gn[3 * lo + 1], gn[3 * lo + 2]);
    o = v3(cen[3 * lo], cen[3 * lo + 1], cen[3 * lo + 2]) + 1.0e-3f * n;
    e.n = n;
    e.r1 = petch::u24(u[2]);
    e.r2 = petch::u24(u[3]);
    e.mode = kLobe;
    R.w = tot > 0.0 ? 1.f : 0.f;
    R.bounce = 0;
    R.ray = ray;
  }
  __device__ __forceinline__ bool hit(Ray& R, const Hit& h, Vec3& o, Vec3 d, Emit& e) const {
    return rr_bounce(R.w, R.bounce, seed, R.ray, bv, h, fl, fixed, S, o, d, e);
  }
};

// _trace3d_cov (threed.py:511-546), the capped coverage trace of cov_transport = 'cap': every
// hit deposits 1; the ray sticks with probability S = bare[face] * beta (word 0 of its bounce
// block against S) and stops at the latest after n_reemit re-emissions; otherwise cosine
// re-emission as in the RR walk. The reference launches it with no boundary handling, so the
// box never wraps. Streams as NeutralRR's: source from block 0, bounce k from block (k, ray, 1).
struct NeutralCap {
  static constexpr bool kReplay = false;
  BvhView bv;
  Box b;
  Source src;
  const float* __restrict__ bare;
  float beta;
  int n_reemit;
  float* fl;
  uint64_t seed;
  int fixed;
  struct Ray {
    int bounce;
    unsigned ray;
  };
  __device__ __forceinline__ const BvhView& view(const Ray&) const { return bv; }
  __device__ __forceinline__ const Box& box(const Ray&) const { return b; }
  __device__ __forceinline__ void start(Ray& R, unsigned ray, Vec3& o, Emit& e) const {
    neutral_source(seed, ray, b, src, o, e);
    R.bounce = 0;
    R.ray = ray;
  }
  __device__ __forceinline__ bool hit(Ray& R, const Hit& h, Vec3& o, Vec3 d, Emit& e) const {
    Vec3 nn = hit_normal(bv, h.slot);
    if (petch::dot(nn, d) > 0.f) nn = -nn;
    deposit(fl, h.face, 1.f, fixed);
    uint32_t u[4];
    petch::philox_block(seed, R.ray, (uint32_t)R.bounce + 1u, 1u, u);
    if (petch::u24(u[0]) < bare[h.face] * beta || R.bounce >= n_reemit) return false;
    ++R.bounce;
    rr_emit(h, nn, u, o, d, e);
    return true;
  }
};

// Path-replay adjoint of the trace (no reference counterpart): for gfl = d loss / d fl it
// accumulates d loss / d bare into gbare, with every ray following the path the trace drew (same
// Philox stream, same roulette and termination decisions, same geometry). Along a ray the
// deposits are w_k at hit k and w_{k+1} = c_k w_k (1 - S_k) with S_k = bare[f_k] beta, where
// c_k = 1, or 1 / p for a ray that survived a roulette with probability p = w / 0.3 (the trace
// gives survivors 0.3 = w / p). Holding p fixed while w varies makes the derivative carry on
// through the roulette, and its expectation is the derivative of the expected flux
// (E[1{u < p} w / p] = w); cutting it at the roulette instead would miss that term (about a
// fifth of the gradient on a trench, measured against common-random-number differences). So for
// every later hit j, d w_j / d S_i = -w_j / (1 - S_i), and
//   d loss / d S_i = -(sum over later hits k of gfl[f_k] w_k) / (1 - S_i) = -(T - P_i) / (1 - S_i),
// with T the path's total of gfl[f_k] w_k and P_i its running sum up to hit i. T is only known
// when the ray ends, so phase 0 accumulates it, then the ray restarts from its source emission
// (origin, lobe, weight and bounce index kept in registers; the Philox stream is keyed by
// (ray, bounce), so the replay redraws the same numbers) and phase 1 replays the path hit for
// hit, adding -beta (T - P_i) / (1 - S_i) to gbare[f_i]. One-hit paths owe nothing and are not
// replayed. T and P are float64, so T - P_i keeps its digits when S_i is close to 1. Most paths
// are short, so phase 0 also keeps the face ids of the first kSegBuf hits in LDS; a path that
// fits is scattered from there when it ends (its weights recomputed from the first one with the
// same steps, its S and gfl read again), and only longer ones are traced again. Face ids alone
// keep the buffer at 16 KB per block.
constexpr int kSegBuf = 16;

// This lane's path buffer: entry k of lane t at [k * kBS + t] (consecutive lanes, consecutive
// banks).
__device__ __forceinline__ int* seg_buf() {
  __shared__ int buf_f[kSegBuf * kBS];
  return buf_f + threadIdx.x;
}

// Like the trace, one launch can carry the two species of a fixed-point iteration: queue ids
// [0, n0) are species 0, the rest species 1 with ids from n0, each with its own beta, stream
// and gfl; both add into the one gbare.
struct NeutralRRAdjoint {
  static constexpr bool kReplay = true;
  BvhView bv;
  Box b;
  Source src;
  const float* __restrict__ bare;
  const float* gfl_sp[2];
  float* gbare;
  float beta_sp[2];
  uint64_t seed_sp[2];
  unsigned n0;
  struct Ray {
    float w, w0;
    int bounce, b0, bend;  // hit index now, at the path's start, of the path's last hit
    int phase;
    int nb;                // hits of the path kept in the LDS buffer, -1 once it overflowed
    int sp;                // species
    unsigned ray;          // id within the species' stream
    double T, P;
    Vec3 o0;
    Emit e0;
  };
  __device__ __forceinline__ const BvhView& view(const Ray&) const { return bv; }
  __device__ __forceinline__ const Box& box(const Ray&) const { return b; }
  __device__ __forceinline__ static void begin_path(Ray& R, const Vec3& o, const Emit& e) {
    R.o0 = o;
    R.e0 = e;
    R.w0 = R.w;
    R.b0 = R.bounce;
    R.phase = 0;
    R.nb = 0;
    R.T = 0.0;
  }
  // the gradients of a buffered path (its last hit owes nothing); every hit but the last
  // continued the ray, so a weight that fell below 0.1 there survived its roulette (0.3)
  __device__ __forceinline__ void scatter(const Ray& R) const {
    const int* B = seg_buf();
    const float beta = beta_sp[R.sp];
    const float* __restrict__ gfl = gfl_sp[R.sp];
    double P = 0.0;
    float w = R.w0;
    for (int k = 0; k < R.nb - 1; ++k) {
      const int f = B[k * kBS];
      const float S = bare[f] * beta;
      P += (double)(gfl[f] * w);
      const double q = 1.0 - (double)S;
      if (q > 0.0) deposit(gbare, f, (float)(-(double)beta * (R.T - P) / q));
      float w2 = w - w * S;
      if (w2 < 0.1f) w2 = 0.3f;
      w = w2;
    }
  }
  __device__ __forceinline__ static bool replay(Ray& R, int bend, Vec3& o, Emit& e) {
    R.bend = bend;
    R.phase = 1;
    R.P = 0.0;
    R.w = R.w0;
    R.bounce = R.b0;
    o = R.o0;
    e = R.e0;
    return true;
  }
  __device__ __forceinline__ void start(Ray& R, unsigned ray, Vec3& o, Emit& e) const {
    R.sp = ray >= n0;
    R.ray = R.sp ? ray - n0 : ray;
    neutral_source(seed_sp[R.sp], R.ray, b, src, o, e);
    R.w = 1.f;
    R.bounce = 0;
    begin_path(R, o, e);
  }
  __device__ __forceinline__ bool hit(Ray& R, const Hit& h, Vec3& o, Vec3 d, Emit& e) const {
    Vec3 nn = hit_normal(bv, h.slot);
    if (petch::dot(nn, d) > 0.f) nn = -nn;
    const int i = R.bounce;
    const float beta = beta_sp[R.sp];
    const float S = bare[h.face] * beta;
    const float gwf = gfl_sp[R.sp][h.face] * R.w;
    const double gw = (double)gwf;
    float w2;
    bool reset;
    uint32_t u[4];
    const bool cont = rr_weight(R.w, S, seed_sp[R.sp], R.ray, R.bounce, w2, reset, u);
    if (R.phase == 0) {
      R.T += gw;
      if (R.nb >= 0 && R.nb < kSegBuf) {
        seg_buf()[R.nb * kBS] = h.face;
        ++R.nb;
      } else {
        R.nb = -1;
      }
      if (!cont && i > R.b0) {
        if (R.nb < 0) return replay(R, i, o, e);
        scatter(R);
      }
    } else {
      R.P += gw;
      const double q = 1.0 - (double)S;
      if (i != R.bend && q > 0.0)
        deposit(gbare, h.face, (float)(-(double)beta * (R.T - R.P) / q));
    }
    if (!cont) return false;
    R.w = w2;
    rr_emit(h, nn, u, o, d, e);
    return true;
  }
  // The ray left the domain after hit R.bounce - 1: a phase-0 path of two or more hits is
  // scattered from the buffer, or replayed if it did not fit; a replayed path (phase 1) ends.
  __device__ __forceinline__ bool miss(Ray& R, Vec3& o, Emit& e) const {
    if (R.phase == 0 && R.bounce - 1 > R.b0) {
      if (R.nb < 0) return replay(R, R.bounce - 1, o, e);
      scatter(R);
    }
    return false;
  }
};

struct IonParams {
  float meanE, sigmaE, A_sp, Eth_sp, B_sp, A_ie, Eth_ie, A_p, Eth_p;
  float inflect, minang, n_l, A_en, minE, thrmin, thrmax;
};

// _trace3d_ion_yield (threed.py:599-712): energy ~ N(meanE, sigmaE) clamped at 0; every hit
// deposits the sputter / ion-enhanced / passivation yields, the arrival weight and the weighted
// cosine; sticking ramps from 1 below thetaRMin to 0 at 90 deg; energy loss by the inflection
// law; coned-cosine re-emission about the specular direction.
// One faithful-ion bounce (threed.py:625-712) of an ion with energy Ein and weight win that hit
// face h: deposits the five yields, then sticks, or loses energy and re-emits about the
// specular direction (returns true with the next ray in o, d). The constants come from F:
// F.P (IonParams), F.sE (square roots of the three threshold energies) and F.out (sp, ie, ip,
// arr, ang accumulators).
template <class Fc>
__device__ __forceinline__ bool ion_bounce(Rng& r, float& Ein, float& win, int& bounce,
                                           const BvhView& bv, const Hit& h, const Fc& F, Vec3& o,
                                           Vec3& d, float E0 = 0.f) {
  const IonParams& P = F.P;
  Vec3 nn = hit_normal(bv, h.slot);
  if (petch::dot(nn, d) > 0.f) nn = -nn;
  const float E = Ein, w = win;
  const float cosT = fminf(fmaxf(-petch::dot(d, nn), 0.f), 1.f);
  const float incAng = acosf(cosT);
  const float sqrtE = sqrtf(E);
  const float f_sp = fmaxf((1.f + P.B_sp * (1.f - cosT * cosT)) * cosT, 0.f);
  float f_ie = 1.f;
  if (cosT < 0.5f) f_ie = fmaxf(3.f - 6.f * incAng / kPi, 0.f);
  if constexpr (Fc::kTangent) {
    // d / d meanE of the three yield deposits: every energy step below scales E by a factor
    // that does not depend on E (its acceptance test compares the factor with [0, 1]), so
    // dE / dE0 = E / E0 along the path, and dE0 / dmeanE = 1 (0 where E0 was clamped to 0,
    // where sqrtE = 0 deposits nothing).
    const float s = w * sqrtE / (2.f * E0);
    deposit(F.out[0], h.face, sqrtE > F.sE[0] ? P.A_sp * f_sp * s : 0.f);
    deposit(F.out[1], h.face, sqrtE > F.sE[1] ? P.A_ie * f_ie * s : 0.f);
    deposit(F.out[2], h.face, sqrtE > F.sE[2] ? P.A_p * f_ie * s : 0.f);
  } else {
    deposit(F.out[0], h.face, P.A_sp * fmaxf(sqrtE - F.sE[0], 0.f) * f_sp * w, F.fixed);
    deposit(F.out[1], h.face, P.A_ie * fmaxf(sqrtE - F.sE[1], 0.f) * f_ie * w, F.fixed);
    deposit(F.out[2], h.face, P.A_p * fmaxf(sqrtE - F.sE[2], 0.f) * f_ie * w, F.fixed);
    deposit(F.out[3], h.face, w, F.fixed);
    deposit(F.out[4], h.face, cosT * w, F.fixed);
  }
  float stick = 1.f;
  if (incAng > P.thrmin) {
    float rr = (incAng - P.thrmin) / (P.thrmax - P.thrmin);
    rr = fminf(fmaxf(rr, 0.f), 1.f);
    stick = 1.f - rr;
  }
  if (stick >= 1.f) return false;
  float Eref;
  if (incAng >= P.inflect)
    Eref = 1.f - (1.f - P.A_en) * (kPi2 - incAng) / (kPi2 - P.inflect);
  else
    Eref = P.A_en * powf(incAng / P.inflect, P.n_l);
  float newE = Eref * E;
  for (int e = 0; e < 4; ++e) {
    float cand = Eref * E + 0.1f * E * r.normal();
    if (cand >= 0.f && cand <= E) {
      newE = cand;
      break;
    }
  }
  newE = fminf(fmaxf(newE, 0.f), E);
  if (newE <= P.minE) return false;
  Ein = newE;
  const float w2 = w - w * stick;
  if (w2 <= 1.0e-4f || ++bounce >= 1024) return false;
  win = w2;
  const Vec3 p = o + h.t * d;
  const Vec3 axis = petch::normalize(d - (2.f * petch::dot(d, nn)) * nn);
  Vec3 tt, bb;
  if (axis.z < -0.9999999f) {
    tt = v3(0.f, -1.f, 0.f);
    bb = v3(-1.f, 0.f, 0.f);
  } else {
    float aa = 1.f / (1.f + axis.z);
    float cc = -axis.x * axis.y * aa;
    tt = v3(1.f - axis.x * axis.x * aa, cc, -axis.x);
    bb = v3(cc, 1.f - axis.y * axis.y * aa, -axis.y);
  }
  const float maxcone = kPi2 - fminf(incAng, P.minang);
  float theta = 0.f;
  for (int t = 0; t < 64; ++t) {
    float uu = sqrtf(r.uniform());
    float ss = sqrtf(fmaxf(1.f - uu, 0.f));
    theta = maxcone * ss;
    if (r.uniform() * theta * uu <= cosf(kPi2 * ss) * sinf(theta)) break;
  }
  const float phi = kTwoPi * r.uniform();
  Vec3 nd = petch::normalize(sinf(theta) * (cosf(phi) * tt + sinf(phi) * bb) + cosf(theta) * axis);
  const float dn = petch::dot(nd, nn);
  if (dn <= 0.f) nd = nd - (2.f * dn) * nn;
  d = nd;
  o = p + 1.0e-4f * nn;
  return true;
}

struct IonYield {
  static constexpr bool kReplay = false;
  static constexpr bool kTangent = false;
  BvhView bv;
  Box b;
  Source src;
  IonParams P;
  float sE[3];
  float* out[5];
  uint64_t seed;
  int fixed;  // out are int64 fixed-point accumulators (deterministic mode)
  struct Ray {
    float E, w;
    int bounce;
    Rng rng;
  };
  __device__ __forceinline__ const BvhView& view(const Ray&) const { return bv; }
  __device__ __forceinline__ const Box& box(const Ray&) const { return b; }
  __device__ __forceinline__ void start(Ray& R, unsigned ray, Vec3& o, Emit& e) const {
    R.rng.init(seed, ray);
    gen_source(R.rng, b, src, o, e.n);
    e.mode = kDirect;
    float E = P.meanE + P.sigmaE * R.rng.normal();
    R.E = E < 0.f ? 0.f : E;
    R.w = 1.f;
    R.bounce = 0;
  }
  __device__ __forceinline__ bool hit(Ray& R, const Hit& h, Vec3& o, Vec3 d, Emit& e) const {
    e.mode = kDirect;
    e.n = d;
    return ion_bounce(R.rng, R.E, R.w, R.bounce, bv, h, *this, o, e.n);
  }
};

// Forward-mode derivative of the faithful-ion trace in the mean ion energy (no reference
// counterpart): the same rays, energies, decisions and paths as IonYield with the same seed;
// out[0..2] receive d (sp, ie, ip) / d meanE per face. The arrival weight and the cosine
// channel do not depend on the energy.
struct IonYieldTangent {
  static constexpr bool kReplay = false;
  static constexpr bool kTangent = true;
  BvhView bv;
  Box b;
  Source src;
  IonParams P;
  float sE[3];
  float* out[3];
  uint64_t seed;
  struct Ray {
    float E, w, E0;
    int bounce;
    Rng rng;
  };
  __device__ __forceinline__ const BvhView& view(const Ray&) const { return bv; }
  __device__ __forceinline__ const Box& box(const Ray&) const { return b; }
  __device__ __forceinline__ void start(Ray& R, unsigned ray, Vec3& o, Emit& e) const {
    R.rng.init(seed, ray);
    gen_source(R.rng, b, src, o, e.n);
    e.mode = kDirect;
    float E = P.meanE + P.sigmaE * R.rng.normal();
    R.E = E < 0.f ? 0.f : E;
    R.E0 = R.E;
    R.w = 1.f;
    R.bounce = 0;
  }
  __device__ __forceinline__ bool hit(Ray& R, const Hit& h, Vec3& o, Vec3 d, Emit& e) const {
    e.mode = kDirect;
    e.n = d;
    return ion_bounce(R.rng, R.E, R.w, R.bounce, bv, h, *this, o, e.n, R.E0);
  }
};

// ---------------------------------------------------------------- batches of features
// A batched launch traces many features (each its own surface, box, coverage and recipe) from
// one ray queue: feature f owns queue ids [end[f-1], end[f]). A lane keeps its ray's feature
// index and the feature's BVH view and box in registers; the per-hit constants are read from
// the feature's descriptor (a few hundred bytes, cache resident).
struct NeutralFeature {
  BvhView bv;
  Box b;
  const float* bare;
  float beta[2];
  float* fl[2];
  uint64_t seed[2];
  unsigned n0;   // rays of species 0; the feature's remaining rays are species 1
  unsigned end;  // exclusive end of the feature's ids in the batch queue
  int fixed;     // fl are int64 fixed-point accumulators (deterministic mode)
  float* dst[2]; // with fixed: the float outputs the accumulators are converted into
};

struct IonFeature {
  static constexpr bool kTangent = false;
  BvhView bv;
  Box b;
  Source src;
  IonParams P;
  float sE[3];
  float* out[5];
  uint64_t seed;
  unsigned end;
  int fixed;      // out are int64 fixed-point accumulators (deterministic mode)
  float* dst[5];  // with fixed: the float outputs the accumulators are converted into
};

// first feature whose id range ends after `ray`
template <class Fd>
__device__ __forceinline__ int locate_feature(const Fd* __restrict__ feat, int nfeat,
                                              unsigned ray) {
  int lo = 0, hi = nfeat - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (feat[mid].end > ray) hi = mid; else lo = mid + 1;
  }
  return lo;
}

struct NeutralRRBatch {
  static constexpr bool kReplay = false;
  const NeutralFeature* __restrict__ feat;
  int nfeat;
  Source src;
  struct Ray {
    float w;
    int bounce;
    int sp;
    int f;
    unsigned ray;  // id within the species' stream
    BvhView bv;
    Box b;
  };
  __device__ __forceinline__ const BvhView& view(const Ray& R) const { return R.bv; }
  __device__ __forceinline__ const Box& box(const Ray& R) const { return R.b; }
  __device__ __forceinline__ void start(Ray& R, unsigned ray, Vec3& o, Emit& e) const {
    R.f = locate_feature(feat, nfeat, ray);
    const NeutralFeature& F = feat[R.f];
    const unsigned local = ray - (R.f ? feat[R.f - 1].end : 0u);
    R.sp = local >= F.n0;
    R.ray = R.sp ? local - F.n0 : local;
    R.bv = F.bv;
    R.b = F.b;
    neutral_source(R.sp ? F.seed[1] : F.seed[0], R.ray, R.b, src, o, e);
    R.w = 1.f;
    R.bounce = 0;
  }
  __device__ __forceinline__ bool hit(Ray& R, const Hit& h, Vec3& o, Vec3 d, Emit& e) const {
    const NeutralFeature& F = feat[R.f];
    const float S = F.bare[h.face] * (R.sp ? F.beta[1] : F.beta[0]);
    return rr_bounce(R.w, R.bounce, R.sp ? F.seed[1] : F.seed[0], R.ray, R.bv, h,
                     R.sp ? F.fl[1] : F.fl[0], F.fixed, S, o, d, e);
  }
};

struct IonYieldBatch {
  static constexpr bool kReplay = false;
  const IonFeature* __restrict__ feat;
  int nfeat;
  struct Ray {
    float E, w;
    int bounce;
    int f;
    Rng rng;
    BvhView bv;
    Box b;
  };
  __device__ __forceinline__ const BvhView& view(const Ray& R) const { return R.bv; }
  __device__ __forceinline__ const Box& box(const Ray& R) const { return R.b; }
  __device__ __forceinline__ void start(Ray& R, unsigned ray, Vec3& o, Emit& e) const {
    R.f = locate_feature(feat, nfeat, ray);
    const IonFeature& F = feat[R.f];
    R.rng.init(F.seed, ray - (R.f ? feat[R.f - 1].end : 0u));
    R.bv = F.bv;
    R.b = F.b;
    gen_source(R.rng, R.b, F.src, o, e.n);
    e.mode = kDirect;
    float E = F.P.meanE + F.P.sigmaE * R.rng.normal();
    R.E = E < 0.f ? 0.f : E;
    R.w = 1.f;
    R.bounce = 0;
  }
  __device__ __forceinline__ bool hit(Ray& R, const Hit& h, Vec3& o, Vec3 d, Emit& e) const {
    e.mode = kDirect;
    e.n = d;
    return ion_bounce(R.rng, R.E, R.w, R.bounce, R.bv, h, feat[R.f], o, e.n);
  }
};

// zero the per-face outputs (float, or int64 with fixed) of every feature of a batch
// (blockIdx.y = feature)
__global__ void k_zero_neutral_batch(const NeutralFeature* __restrict__ feat) {
  const NeutralFeature& F = feat[blockIdx.y];
  for (int i = blockIdx.x * kBS + threadIdx.x; i < F.bv.nf; i += gridDim.x * kBS) {
    for (int k = 0; k < 2; ++k) {
      if (F.fixed) ((unsigned long long*)F.fl[k])[i] = 0ull;
      else F.fl[k][i] = 0.f;
    }
  }
}

__global__ void k_zero_ion_batch(const IonFeature* __restrict__ feat) {
  const IonFeature& F = feat[blockIdx.y];
  for (int i = blockIdx.x * kBS + threadIdx.x; i < F.bv.nf; i += gridDim.x * kBS) {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      if (F.fixed) ((unsigned long long*)F.out[k])[i] = 0ull;
      else F.out[k][i] = 0.f;
    }
  }
}

__device__ __forceinline__ float fix_to_float(unsigned long long q) {
  return (float)((double)(long long)q * (1.0 / kFixScale));
}

// the deterministic mode's last pass: int64 fixed-point sums -> the float outputs
__global__ void k_fix_finish(const unsigned long long* __restrict__ q, float* __restrict__ out,
                             int n) {
  const int i = blockIdx.x * kBS + threadIdx.x;
  if (i < n) out[i] = fix_to_float(q[i]);
}

__global__ void k_fix_finish_neutral_batch(const NeutralFeature* __restrict__ feat) {
  const NeutralFeature& F = feat[blockIdx.y];
  for (int i = blockIdx.x * kBS + threadIdx.x; i < F.bv.nf; i += gridDim.x * kBS)
    for (int k = 0; k < 2; ++k) F.dst[k][i] = fix_to_float(((const unsigned long long*)F.fl[k])[i]);
}

__global__ void k_fix_finish_ion_batch(const IonFeature* __restrict__ feat) {
  const IonFeature& F = feat[blockIdx.y];
  for (int i = blockIdx.x * kBS + threadIdx.x; i < F.bv.nf; i += gridDim.x * kBS)
#pragma unroll
    for (int k = 0; k < 5; ++k)
      F.dst[k][i] = fix_to_float(((const unsigned long long*)F.out[k])[i]);
}

// The neutral loop fits 96 registers without spills, which gives 5 waves per SIMD; the ion
// loop needs more and runs best at the compiler's own choice (4 waves per SIMD), measured.
__global__ void __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(5)))
k_trace_neutral_rr(NeutralRR pol, RayQueue q, unsigned long long* __restrict__ hits) {
  persistent_trace(pol, q, hits);
}

__global__ void __launch_bounds__(kBS)
k_trace_neutral_rr_adjoint(NeutralRRAdjoint pol, RayQueue q, unsigned long long* __restrict__ hits) {
  persistent_trace(pol, q, hits);
}

__global__ void __launch_bounds__(kBS)
k_trace_ion_yield(IonYield pol, RayQueue q, unsigned long long* __restrict__ hits) {
  persistent_trace(pol, q, hits);
}

__global__ void __launch_bounds__(kBS)
k_trace_ion_yield_tangent(IonYieldTangent pol, RayQueue q, unsigned long long* __restrict__ hits) {
  persistent_trace(pol, q, hits);
}

__global__ void __launch_bounds__(kBS)
k_trace_neutral_rr_batch(NeutralRRBatch pol, RayQueue q, unsigned long long* __restrict__ hits) {
  persistent_trace(pol, q, hits);
}

__global__ void __launch_bounds__(kBS)
k_trace_ion_yield_batch(IonYieldBatch pol, RayQueue q, unsigned long long* __restrict__ hits) {
  persistent_trace(pol, q, hits);
}

__global__ void __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(5)))
k_trace_redep(Redep pol, RayQueue q, unsigned long long* __restrict__ hits) {
  persistent_trace(pol, q, hits);
}

__global__ void __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(5)))
k_trace_neutral_cap(NeutralCap pol, RayQueue q, unsigned long long* __restrict__ hits) {
  persistent_trace(pol, q, hits);
}

// Emission weight of a face for redeposition: max(rate * area, 0), a non-finite product as 0.
__device__ __forceinline__ double redep_weight(const float* __restrict__ rate,
                                               const float* __restrict__ area, int f) {
  const double w = (double)rate[f] * (double)area[f];
  return (w > 0.0 && w <= 1.0e300) ? w : 0.0;
}

// Inclusive and exclusive prefix of v over a workgroup of BS threads (BS a multiple of 64) in
// float64; the workgroup total goes to `total`.
template <int BS>
__device__ __forceinline__ double block_scan_f64(double v, double& excl, double& total) {
  __shared__ double wsum[BS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  double ex = __shfl_up(inc, 1, 64);
  if (lane == 0) ex = 0.0;
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  double base = 0.0, tot = 0.0;
#pragma unroll
  for (int i = 0; i < BS / 64; ++i) {
    const double s = wsum[i];
    base += (i < w) ? s : 0.0;
    tot += s;
  }
  __syncthreads();
  total = tot;
  excl = base + ex;
  return base + inc;
}

// Redeposition CDF, pass 1: the total emission weight of every workgroup's kBS faces.
__global__ void __launch_bounds__(kBS) k_redep_partials(const float* __restrict__ rate,
                                                        const float* __restrict__ area, int nf,
                                                        double* __restrict__ part) {
  const int f = blockIdx.x * kBS + threadIdx.x;
  double ex, tot;
  block_scan_f64<kBS>(f < nf ? redep_weight(rate, area, f) : 0.0, ex, tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// Pass 2: exclusive prefix of the nb workgroup totals, in place, by one workgroup.
__global__ void __launch_bounds__(1024) k_redep_scan_partials(double* __restrict__ part, int nb) {
  double carry = 0.0;
  for (int c0 = 0; c0 < nb; c0 += 1024) {
    const int i = c0 + threadIdx.x;
    double ex, tot;
    block_scan_f64<1024>(i < nb ? part[i] : 0.0, ex, tot);
    if (i < nb) part[i] = carry + ex;
    carry += tot;
  }
}

// Pass 3: cum[f] = the workgroup's offset + the inclusive prefix within it.
__global__ void __launch_bounds__(kBS) k_redep_cum(const float* __restrict__ rate,
                                                   const float* __restrict__ area, int nf,
                                                   const double* __restrict__ part,
                                                   double* __restrict__ cum) {
  const int f = blockIdx.x * kBS + threadIdx.x;
  double ex, tot;
  const double inc = block_scan_f64<kBS>(f < nf ? redep_weight(rate, area, f) : 0.0, ex, tot);
  if (f < nf) cum[f] = part[blockIdx.x] + inc;
}

// The redeposited flux density (threed.py:1605): out = fl * (total / n) / max(area, 1e-9), or 0
// when nothing was emitted; fl is the trace's per-face deposit, read from out.
__global__ void k_redep_finish(float* __restrict__ out, const float* __restrict__ area,
                               const double* __restrict__ cum, int nf, int n) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double tot = cum[nf - 1];
  const double a = fmax((double)area[f], 1.0e-9);
  out[f] = tot > 0.0 ? (float)((double)out[f] * (tot / (double)n) / a) : 0.f;
}

__global__ void __launch_bounds__(kBS) k_trace_legacy(BvhView bv, Source src, Box b,
                                                      float sticking, int n_reemit, int specular,
                                                      float cos_thr, float eta, int n,
                                                      uint64_t seed, float* __restrict__ flux,
                                                      float* __restrict__ angacc) {
  int p = blockIdx.x * kBS + threadIdx.x;
  if (p >= n) return;
  Rng r;
  r.init(seed, (uint32_t)p);
  Vec3 o, d;
  gen_source(r, b, src, o, d);
  float w = 1.f;
  for (int bounce = 0; bounce <= n_reemit; ++bounce) {
    Hit h = petch::bvh_closest_hit(bv, o, d, 1.0e6f);
    if (h.face < 0) break;
    Vec3 nn = hit_normal(bv, h.slot);
    if (petch::dot(nn, d) > 0.f) nn = -nn;
    float cosang = fmaxf(-petch::dot(d, nn), 0.f);
    if (specular) {
      if (cosang >= cos_thr || bounce == n_reemit) {
        deposit(flux, h.face, w);
        deposit(angacc, h.face, w * cosang);
        break;
      }
      float R = (cos_thr - cosang) / cos_thr;
      float dep = w * (1.f - R);
      deposit(flux, h.face, dep);
      deposit(angacc, h.face, dep * cosang);
      w = w * R * eta;
      Vec3 hit = o + h.t * d;
      d = d - (2.f * petch::dot(d, nn)) * nn;
      o = hit + 1.0e-4f * nn;
    } else {
      float u = r.uniform();
      if (u < sticking || bounce == n_reemit) {
        deposit(flux, h.face, w);
        deposit(angacc, h.face, w * cosang);
        break;
      }
      Vec3 hit = o + h.t * d;
      d = cosine_about(r, nn);
      o = hit + 1.0e-4f * nn;
    }
  }
}

// n_ff cosine rays per face from centroid + 1e-3 n (threed.py:949-955); hit face or -1.
__global__ void __launch_bounds__(kBS) k_trace_ff(BvhView bv, Box b,
                                                  const float* __restrict__ cen,
                                                  const float* __restrict__ gn, int nf, int n_ff,
                                                  uint64_t seed, int* __restrict__ hitface) {
  long long p = (long long)blockIdx.x * kBS + threadIdx.x;
  if (p >= (long long)nf * n_ff) return;
  int j = (int)(p / n_ff);
  Rng r;
  r.init(seed, (uint32_t)p);
  Vec3 nn = v3(gn[3 * j], gn[3 * j + 1], gn[3 * j + 2]);
  Vec3 o = v3(cen[3 * j], cen[3 * j + 1], cen[3 * j + 2]) + 1.0e-3f * nn;
  // _cosine_dirs (threed.py:1571-1579): t = cross(a, n) / (|cross(a, n)| + 1e-12)
  Vec3 a = (fabsf(nn.x) > 0.9f) ? v3(0.f, 1.f, 0.f) : v3(1.f, 0.f, 0.f);
  Vec3 c = petch::cross(a, nn);
  Vec3 t = (1.f / (petch::length(c) + 1.0e-12f)) * c;
  Vec3 bb = petch::cross(nn, t);
  float u1 = r.uniform(), u2 = r.uniform();
  float ct = sqrtf(u1), st = sqrtf(1.f - u1), ph = kTwoPi * u2;
  Vec3 d = (st * cosf(ph)) * t + (st * sinf(ph)) * bb + ct * nn;
  Hit h = trace_bc(bv, b, o, d);
  hitface[p] = h.face;
}

__global__ void k_gen_source(Box b, Source src, int n, uint64_t seed, float* __restrict__ origin,
                             float* __restrict__ dir) {
  int p = blockIdx.x * kBS + threadIdx.x;
  if (p >= n) return;
  Rng r;
  r.init(seed, (uint32_t)p);
  Vec3 o, d;
  gen_source(r, b, src, o, d);
  origin[3 * p] = o.x;
  origin[3 * p + 1] = o.y;
  origin[3 * p + 2] = o.z;
  dir[3 * p] = d.x;
  dir[3 * p + 1] = d.y;
  dir[3 * p + 2] = d.z;
}

Box make_box(const petch_box* b) { return Box{b->Lx, b->Ly, b->Lz, b->z_src, b->periodic}; }

Source make_source(const petch_source* s) {
  return Source{s->kind, s->sigma, s->exponent};
}

// Grid of a persistent trace: as many workgroups as can be resident at once (CUs x the
// kernel's occupancy, from its register and LDS use), never more than the rays need.
unsigned persistent_blocks(petch_ctx* ctx, const void* kernel, int n) {
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
          hipSuccess ||
      cus <= 0)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBS, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 4;
  if (ctx->trace_blocks_per_cu > 0 && ctx->trace_blocks_per_cu < per_cu)
    per_cu = ctx->trace_blocks_per_cu;
  long long want = (long long)cus * per_cu;
  long long need = ((long long)n + kBS - 1) / kBS;
  return (unsigned)(need < want ? need : want);
}

int make_queue(petch_ctx* ctx, int n, RayQueue& q) {
  unsigned long long* c = petch_hit_counter(ctx, 2);
  if (!c) return PETCH_E_HIP;
  q.next = (unsigned*)c;
  q.n = n;
  PETCH_HIP(ctx, hipMemsetAsync(q.next, 0, sizeof(unsigned), ctx->stream));
  return PETCH_OK;
}

int zero(petch_ctx* ctx, float* p, int n) {
  if (p && n > 0) PETCH_HIP(ctx, hipMemsetAsync(p, 0, (size_t)n * sizeof(float), ctx->stream));
  return PETCH_OK;
}

// Deterministic mode: k zeroed int64 fixed-point accumulators of n faces each (null on error).
unsigned long long* det_accumulators(petch_ctx* ctx, int k, long long n) {
  const size_t bytes = (size_t)k * (size_t)(n > 0 ? n : 1) * sizeof(unsigned long long);
  auto* q = (unsigned long long*)petch_scratch(ctx, SCR_DET, bytes);
  if (!q) return nullptr;
  if (hipMemsetAsync(q, 0, bytes, ctx->stream) != hipSuccess) {
    petch_set_error(ctx, PETCH_E_HIP, "deterministic mode: clearing the accumulators failed");
    return nullptr;
  }
  return q;
}

int det_finish(petch_ctx* ctx, const unsigned long long* q, float* out, int n) {
  if (n <= 0) return PETCH_OK;
  hipLaunchKernelGGL(k_fix_finish, dim3(petch_blocks(n, kBS)), dim3(kBS), 0, ctx->stream, q, out,
                     n);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

}  // namespace

extern "C" int petch_gen_source(petch_ctx* ctx, const petch_box* box, const petch_source* src,
                                int n, uint64_t seed, float* origin, float* dir) {
  PETCH_ARG(ctx, ctx && box && src && origin && dir, "gen_source: null pointer");
  PETCH_ARG(ctx, n >= 0, "gen_source: n < 0");
  if (n == 0) return PETCH_OK;
  hipLaunchKernelGGL(k_gen_source, dim3(petch_blocks(n, kBS)), dim3(kBS), 0, ctx->stream,
                     make_box(box), make_source(src), n, seed, origin, dir);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

namespace {

int launch_neutral(petch_ctx* ctx, const NeutralRR& pol, int n) {
  RayQueue q;
  int rc = make_queue(ctx, n, q);
  if (rc) return rc;
  const unsigned grid = persistent_blocks(ctx, (const void*)k_trace_neutral_rr, n);
  hipLaunchKernelGGL(k_trace_neutral_rr, dim3(grid), dim3(kBS), 0, ctx->stream, pol, q,
                     petch_hit_counter(ctx, 0));
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

}  // namespace

extern "C" int petch_trace_neutral_rr(petch_ctx* ctx, const petch_bvh* bvh, const petch_box* box,
                                      const petch_source* src, const float* bare, float beta,
                                      int n, uint64_t seed, float* fl) {
  PETCH_ARG(ctx, ctx && bvh && box && src && bare && fl, "trace_neutral_rr: null pointer");
  PETCH_ARG(ctx, n >= 0, "trace_neutral_rr: n < 0");
  int rc = zero(ctx, fl, bvh->nf);
  if (rc) return rc;
  if (n == 0 || bvh->nf == 0) return PETCH_OK;
  NeutralRR pol{petch_bvh_view(bvh), make_box(box), make_source(src), bare, {beta, beta},
                {fl, fl}, {seed, seed}, (unsigned)n};
  if (!ctx->deterministic) return launch_neutral(ctx, pol, n);
  unsigned long long* q = det_accumulators(ctx, 1, bvh->nf);
  if (!q) return PETCH_E_HIP;
  pol.fl[0] = pol.fl[1] = (float*)q;
  pol.fixed = 1;
  rc = launch_neutral(ctx, pol, n);
  if (rc) return rc;
  return det_finish(ctx, q, fl, bvh->nf);
}

extern "C" int petch_trace_neutral_cap(petch_ctx* ctx, const petch_bvh* bvh,
                                       const petch_box* box, const petch_source* src,
                                       const float* bare, float beta, int n_reemit, int n,
                                       uint64_t seed, float* fl) {
  PETCH_ARG(ctx, ctx && bvh && box && src && bare && fl, "trace_neutral_cap: null pointer");
  PETCH_ARG(ctx, n >= 0, "trace_neutral_cap: n < 0");
  PETCH_ARG(ctx, n_reemit >= 0, "trace_neutral_cap: n_reemit < 0");
  int rc = zero(ctx, fl, bvh->nf);
  if (rc) return rc;
  if (n == 0 || bvh->nf == 0) return PETCH_OK;
  Box b = make_box(box);
  b.periodic = 0;  // _trace3d_cov applies no boundary condition
  NeutralCap pol{petch_bvh_view(bvh), b, make_source(src), bare, beta, n_reemit, fl, seed, 0};
  unsigned long long* q = nullptr;
  if (ctx->deterministic) {
    q = det_accumulators(ctx, 1, bvh->nf);
    if (!q) return PETCH_E_HIP;
    pol.fl = (float*)q;
    pol.fixed = 1;
  }
  RayQueue rq;
  rc = make_queue(ctx, n, rq);
  if (rc) return rc;
  const unsigned grid = persistent_blocks(ctx, (const void*)k_trace_neutral_cap, n);
  hipLaunchKernelGGL(k_trace_neutral_cap, dim3(grid), dim3(kBS), 0, ctx->stream, pol, rq,
                     petch_hit_counter(ctx, 0));
  PETCH_LAUNCH_CHECK(ctx);
  return q ? det_finish(ctx, q, fl, bvh->nf) : PETCH_OK;
}

extern "C" int petch_trace_redeposition(petch_ctx* ctx, const petch_bvh* bvh,
                                        const float* centroids, const float* gas_normals,
                                        const float* areas, const float* rate, float s_redep,
                                        int n, uint64_t seed, float* out) {
  PETCH_ARG(ctx, ctx && bvh && centroids && gas_normals && areas && rate && out,
            "trace_redeposition: null pointer");
  PETCH_ARG(ctx, n >= 0, "trace_redeposition: n < 0");
  const int nf = bvh->nf;
  int rc = zero(ctx, out, nf);
  if (rc) return rc;
  if (n == 0 || nf == 0) return PETCH_OK;
  const int nb = (int)petch_blocks(nf, kBS);
  auto* cum = (double*)petch_scratch(ctx, SCR_REDEP, ((size_t)nf + (size_t)nb) * sizeof(double));
  if (!cum) return PETCH_E_HIP;
  double* part = cum + nf;
  hipLaunchKernelGGL(k_redep_partials, dim3(nb), dim3(kBS), 0, ctx->stream, rate, areas, nf,
                     part);
  hipLaunchKernelGGL(k_redep_scan_partials, dim3(1), dim3(1024), 0, ctx->stream, part, nb);
  hipLaunchKernelGGL(k_redep_cum, dim3(nb), dim3(kBS), 0, ctx->stream, rate, areas, nf, part,
                     cum);
  PETCH_LAUNCH_CHECK(ctx);
  // open box: no lateral wrap (threed.py:1604 passes periodic = 0)
  Redep pol{petch_bvh_view(bvh), Box{1.f, 1.f, 1.f, 1.f, 0}, centroids, gas_normals, cum, nf,
            s_redep, out, seed, 0};
  unsigned long long* q = nullptr;
  if (ctx->deterministic) {
    q = det_accumulators(ctx, 1, nf);
    if (!q) return PETCH_E_HIP;
    pol.fl = (float*)q;
    pol.fixed = 1;
  }
  RayQueue rq;
  rc = make_queue(ctx, n, rq);
  if (rc) return rc;
  const unsigned grid = persistent_blocks(ctx, (const void*)k_trace_redep, n);
  hipLaunchKernelGGL(k_trace_redep, dim3(grid), dim3(kBS), 0, ctx->stream, pol, rq,
                     (unsigned long long*)nullptr);
  PETCH_LAUNCH_CHECK(ctx);
  if (q) {
    rc = det_finish(ctx, q, out, nf);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_redep_finish, dim3(petch_blocks(nf, kBS)), dim3(kBS), 0, ctx->stream, out,
                     areas, cum, nf, n);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

namespace {

int launch_neutral_adjoint(petch_ctx* ctx, const NeutralRRAdjoint& pol, int n) {
  RayQueue q;
  int rc = make_queue(ctx, n, q);
  if (rc) return rc;
  const unsigned grid = persistent_blocks(ctx, (const void*)k_trace_neutral_rr_adjoint, n);
  hipLaunchKernelGGL(k_trace_neutral_rr_adjoint, dim3(grid), dim3(kBS), 0, ctx->stream, pol, q,
                     (unsigned long long*)nullptr);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

}  // namespace

extern "C" int petch_trace_neutral_rr_adjoint(petch_ctx* ctx, const petch_bvh* bvh,
                                              const petch_box* box, const petch_source* src,
                                              const float* bare, float beta, int n, uint64_t seed,
                                              const float* gfl, float* gbare) {
  PETCH_ARG(ctx, ctx && bvh && box && src && bare && gfl && gbare,
            "trace_neutral_rr_adjoint: null pointer");
  PETCH_ARG(ctx, n >= 0, "trace_neutral_rr_adjoint: n < 0");
  PETCH_ARG(ctx, gbare != bare && gbare != gfl,
            "trace_neutral_rr_adjoint: gbare must not alias bare or gfl");
  int rc = zero(ctx, gbare, bvh->nf);
  if (rc) return rc;
  if (n == 0 || bvh->nf == 0) return PETCH_OK;
  NeutralRRAdjoint pol{petch_bvh_view(bvh), make_box(box), make_source(src), bare, {gfl, gfl},
                       gbare, {beta, beta}, {seed, seed}, (unsigned)n};
  return launch_neutral_adjoint(ctx, pol, n);
}

extern "C" int petch_trace_neutral_rr_pair_adjoint(petch_ctx* ctx, const petch_bvh* bvh,
                                                   const petch_box* box, const petch_source* src,
                                                   const float* bare, float beta_a, int n_a,
                                                   uint64_t seed_a, const float* gfl_a,
                                                   float beta_b, int n_b, uint64_t seed_b,
                                                   const float* gfl_b, float* gbare) {
  PETCH_ARG(ctx, ctx && bvh && box && src && bare && gfl_a && gfl_b && gbare,
            "trace_neutral_rr_pair_adjoint: null pointer");
  PETCH_ARG(ctx, n_a >= 0 && n_b >= 0, "trace_neutral_rr_pair_adjoint: negative ray count");
  PETCH_ARG(ctx, (long long)n_a + n_b <= 0x7fffffffLL,
            "trace_neutral_rr_pair_adjoint: n_a + n_b > 2^31-1");
  PETCH_ARG(ctx, gbare != bare && gbare != gfl_a && gbare != gfl_b,
            "trace_neutral_rr_pair_adjoint: gbare must not alias bare or the gfl inputs");
  int rc = zero(ctx, gbare, bvh->nf);
  if (rc) return rc;
  const int n = n_a + n_b;
  if (n == 0 || bvh->nf == 0) return PETCH_OK;
  NeutralRRAdjoint pol{petch_bvh_view(bvh), make_box(box), make_source(src), bare, {gfl_a, gfl_b},
                       gbare, {beta_a, beta_b}, {seed_a, seed_b}, (unsigned)n_a};
  return launch_neutral_adjoint(ctx, pol, n);
}

extern "C" int petch_trace_neutral_rr_pair(petch_ctx* ctx, const petch_bvh* bvh,
                                           const petch_box* box, const petch_source* src,
                                           const float* bare, float beta_a, int n_a,
                                           uint64_t seed_a, float* fl_a, float beta_b, int n_b,
                                           uint64_t seed_b, float* fl_b) {
  PETCH_ARG(ctx, ctx && bvh && box && src && bare && fl_a && fl_b,
            "trace_neutral_rr_pair: null pointer");
  PETCH_ARG(ctx, n_a >= 0 && n_b >= 0, "trace_neutral_rr_pair: negative ray count");
  PETCH_ARG(ctx, (long long)n_a + n_b <= 0x7fffffffLL,
            "trace_neutral_rr_pair: n_a + n_b > 2^31-1");
  PETCH_ARG(ctx, fl_a != fl_b, "trace_neutral_rr_pair: the two species need distinct outputs");
  int rc = zero(ctx, fl_a, bvh->nf);
  if (rc) return rc;
  rc = zero(ctx, fl_b, bvh->nf);
  if (rc) return rc;
  const int n = n_a + n_b;
  if (n == 0 || bvh->nf == 0) return PETCH_OK;
  NeutralRR pol{petch_bvh_view(bvh), make_box(box), make_source(src), bare, {beta_a, beta_b},
                {fl_a, fl_b}, {seed_a, seed_b}, (unsigned)n_a};
  if (!ctx->deterministic) return launch_neutral(ctx, pol, n);
  const int nf = bvh->nf;
  unsigned long long* q = det_accumulators(ctx, 2, nf);
  if (!q) return PETCH_E_HIP;
  pol.fl[0] = (float*)q;
  pol.fl[1] = (float*)(q + nf);
  pol.fixed = 1;
  rc = launch_neutral(ctx, pol, n);
  if (!rc) rc = det_finish(ctx, q, fl_a, nf);
  if (!rc) rc = det_finish(ctx, q + nf, fl_b, nf);
  return rc;
}

extern "C" int petch_trace_ion_yield_tangent(petch_ctx* ctx, const petch_bvh* bvh,
                                             const petch_box* box, const petch_source* src,
                                             const petch_ion_params* ip, int n, uint64_t seed,
                                             float* dsp, float* die, float* dip) {
  PETCH_ARG(ctx, ctx && bvh && box && src && ip && dsp && die && dip,
            "trace_ion_yield_tangent: null pointer");
  PETCH_ARG(ctx, n >= 0, "trace_ion_yield_tangent: n < 0");
  float* outs[3] = {dsp, die, dip};
  for (int k = 0; k < 3; ++k) {
    int rc = zero(ctx, outs[k], bvh->nf);
    if (rc) return rc;
  }
  if (n == 0 || bvh->nf == 0) return PETCH_OK;
  IonParams P{ip->meanE,  ip->sigmaE, ip->A_sp,    ip->Eth_sp, ip->B_sp, ip->A_ie,
              ip->Eth_ie, ip->A_p,    ip->Eth_p,   ip->inflect, ip->minang, ip->n_l,
              ip->A_en,   ip->minE,   ip->thrmin,  ip->thrmax};
  RayQueue q;
  int rc2 = make_queue(ctx, n, q);
  if (rc2) return rc2;
  IonYieldTangent pol{petch_bvh_view(bvh), make_box(box), make_source(src), P,
                      {sqrtf(P.Eth_sp), sqrtf(P.Eth_ie), sqrtf(P.Eth_p)}, {dsp, die, dip}, seed};
  const unsigned grid = persistent_blocks(ctx, (const void*)k_trace_ion_yield_tangent, n);
  hipLaunchKernelGGL(k_trace_ion_yield_tangent, dim3(grid), dim3(kBS), 0, ctx->stream, pol, q,
                     (unsigned long long*)nullptr);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_trace_ion_yield(petch_ctx* ctx, const petch_bvh* bvh, const petch_box* box,
                                     const petch_source* src, const petch_ion_params* ip, int n,
                                     uint64_t seed, float* sp, float* ie, float* ipass,
                                     float* arr, float* ang) {
  PETCH_ARG(ctx, ctx && bvh && box && src && ip && sp && ie && ipass && arr && ang,
            "trace_ion_yield: null pointer");
  PETCH_ARG(ctx, n >= 0, "trace_ion_yield: n < 0");
  float* outs[5] = {sp, ie, ipass, arr, ang};
  for (int k = 0; k < 5; ++k) {
    int rc = zero(ctx, outs[k], bvh->nf);
    if (rc) return rc;
  }
  if (n == 0 || bvh->nf == 0) return PETCH_OK;
  IonParams P{ip->meanE,  ip->sigmaE, ip->A_sp,    ip->Eth_sp, ip->B_sp, ip->A_ie,
              ip->Eth_ie, ip->A_p,    ip->Eth_p,   ip->inflect, ip->minang, ip->n_l,
              ip->A_en,   ip->minE,   ip->thrmin,  ip->thrmax};
  RayQueue q;
  int rc2 = make_queue(ctx, n, q);
  if (rc2) return rc2;
  IonYield pol{petch_bvh_view(bvh), make_box(box), make_source(src), P,
               {sqrtf(P.Eth_sp), sqrtf(P.Eth_ie), sqrtf(P.Eth_p)}, {sp, ie, ipass, arr, ang},
               seed};
  const int nf = bvh->nf;
  unsigned long long* acc = nullptr;
  if (ctx->deterministic) {
    acc = det_accumulators(ctx, 5, nf);
    if (!acc) return PETCH_E_HIP;
    for (int k = 0; k < 5; ++k) pol.out[k] = (float*)(acc + (size_t)k * nf);
    pol.fixed = 1;
  }
  const unsigned grid = persistent_blocks(ctx, (const void*)k_trace_ion_yield, n);
  hipLaunchKernelGGL(k_trace_ion_yield, dim3(grid), dim3(kBS), 0, ctx->stream, pol, q,
                     petch_hit_counter(ctx, 1));
  PETCH_LAUNCH_CHECK(ctx);
  if (acc)
    for (int k = 0; k < 5; ++k) {
      const int rc3 = det_finish(ctx, acc + (size_t)k * nf, outs[k], nf);
      if (rc3) return rc3;
    }
  return PETCH_OK;
}

extern "C" int petch_trace_neutral_rr_batch(petch_ctx* ctx, const petch_source* src, int nfeat,
                                            const petch_neutral_feature* feats) {
  PETCH_ARG(ctx, ctx && src && (nfeat == 0 || feats), "trace_neutral_rr_batch: null pointer");
  PETCH_ARG(ctx, nfeat >= 0, "trace_neutral_rr_batch: nfeat < 0");
  if (nfeat == 0) return PETCH_OK;
  std::vector<NeutralFeature> t(nfeat);
  long long end = 0;
  int max_nf = 0;
  for (int f = 0; f < nfeat; ++f) {
    const petch_neutral_feature& a = feats[f];
    PETCH_ARG(ctx, a.bvh, "trace_neutral_rr_batch: feature %d has no BVH", f);
    // an empty surface has empty outputs (possibly NULL) and takes no rays
    PETCH_ARG(ctx, a.bvh->nf == 0 || (a.bare && a.fl_a && a.fl_b),
              "trace_neutral_rr_batch: feature %d has a null pointer", f);
    PETCH_ARG(ctx, a.bvh->nf == 0 || a.fl_a != a.fl_b,
              "trace_neutral_rr_batch: feature %d: outputs must differ", f);
    PETCH_ARG(ctx, a.n_a >= 0 && a.n_b >= 0, "trace_neutral_rr_batch: feature %d: n < 0", f);
    NeutralFeature& d = t[f];
    d.bv = petch_bvh_view(a.bvh);
    d.b = make_box(&a.box);
    d.bare = a.bare;
    d.beta[0] = a.beta_a;
    d.beta[1] = a.beta_b;
    d.fl[0] = a.fl_a;
    d.fl[1] = a.fl_b;
    d.seed[0] = a.seed_a;
    d.seed[1] = a.seed_b;
    d.fixed = 0;
    d.dst[0] = a.fl_a;
    d.dst[1] = a.fl_b;
    const long long n = a.bvh->nf > 0 ? (long long)a.n_a + a.n_b : 0;
    d.n0 = a.bvh->nf > 0 ? (unsigned)a.n_a : 0u;
    end += n;
    PETCH_ARG(ctx, end <= 0x7fffffffLL, "trace_neutral_rr_batch: more than 2^31-1 rays");
    d.end = (unsigned)end;
    if (a.bvh->nf > max_nf) max_nf = a.bvh->nf;
  }
  if (ctx->deterministic) {
    long long total = 0;
    for (int f = 0; f < nfeat; ++f) total += 2LL * feats[f].bvh->nf;
    unsigned long long* acc = det_accumulators(ctx, 1, total);
    if (!acc) return PETCH_E_HIP;
    long long off = 0;
    for (int f = 0; f < nfeat; ++f) {
      const int nf = feats[f].bvh->nf;
      t[f].fl[0] = (float*)(acc + off);
      t[f].fl[1] = (float*)(acc + off + nf);
      t[f].fixed = 1;
      off += 2LL * nf;
    }
  }
  const NeutralFeature* dev =
      (const NeutralFeature*)petch_upload_table(ctx, t.data(), t.size() * sizeof(NeutralFeature));
  if (!dev) return PETCH_E_HIP;
  const unsigned gx = petch_blocks(max_nf, kBS) < 64 ? petch_blocks(max_nf, kBS) : 64;
  if (max_nf > 0 && !ctx->deterministic) {
    hipLaunchKernelGGL(k_zero_neutral_batch, dim3(gx, nfeat), dim3(kBS), 0, ctx->stream, dev);
    PETCH_LAUNCH_CHECK(ctx);
  }
  if (end > 0) {
    RayQueue q;
    int rc = make_queue(ctx, (int)end, q);
    if (rc) return rc;
    NeutralRRBatch pol{dev, nfeat, make_source(src)};
    const unsigned grid =
        persistent_blocks(ctx, (const void*)k_trace_neutral_rr_batch, (int)end);
    hipLaunchKernelGGL(k_trace_neutral_rr_batch, dim3(grid), dim3(kBS), 0, ctx->stream, pol, q,
                       petch_hit_counter(ctx, 0));
    PETCH_LAUNCH_CHECK(ctx);
  }
  if (max_nf > 0 && ctx->deterministic) {
    hipLaunchKernelGGL(k_fix_finish_neutral_batch, dim3(gx, nfeat), dim3(kBS), 0, ctx->stream,
                       dev);
    PETCH_LAUNCH_CHECK(ctx);
  }
  return PETCH_OK;
}

extern "C" int petch_trace_ion_yield_batch(petch_ctx* ctx, int nfeat,
                                           const petch_ion_feature* feats) {
  PETCH_ARG(ctx, ctx && (nfeat == 0 || feats), "trace_ion_yield_batch: null pointer");
  PETCH_ARG(ctx, nfeat >= 0, "trace_ion_yield_batch: nfeat < 0");
  if (nfeat == 0) return PETCH_OK;
  std::vector<IonFeature> t(nfeat);
  long long end = 0;
  int max_nf = 0;
  for (int f = 0; f < nfeat; ++f) {
    const petch_ion_feature& a = feats[f];
    PETCH_ARG(ctx, a.bvh, "trace_ion_yield_batch: feature %d has no BVH", f);
    PETCH_ARG(ctx, a.bvh->nf == 0 || (a.sp && a.ie && a.ipass && a.arr && a.ang),
              "trace_ion_yield_batch: feature %d has a null pointer", f);
    PETCH_ARG(ctx, a.n >= 0, "trace_ion_yield_batch: feature %d: n < 0", f);
    IonFeature& d = t[f];
    d.bv = petch_bvh_view(a.bvh);
    d.b = make_box(&a.box);
    d.src = make_source(&a.src);
    const petch_ion_params& ip = a.ip;
    d.P = IonParams{ip.meanE,  ip.sigmaE, ip.A_sp,   ip.Eth_sp,  ip.B_sp,   ip.A_ie,
                    ip.Eth_ie, ip.A_p,    ip.Eth_p,  ip.inflect, ip.minang, ip.n_l,
                    ip.A_en,   ip.minE,   ip.thrmin, ip.thrmax};
    d.sE[0] = sqrtf(ip.Eth_sp);
    d.sE[1] = sqrtf(ip.Eth_ie);
    d.sE[2] = sqrtf(ip.Eth_p);
    d.out[0] = a.sp;
    d.out[1] = a.ie;
    d.out[2] = a.ipass;
    d.out[3] = a.arr;
    d.out[4] = a.ang;
    for (int k = 0; k < 5; ++k) d.dst[k] = d.out[k];
    d.fixed = 0;
    d.seed = a.seed;
    end += a.bvh->nf > 0 ? a.n : 0;
    PETCH_ARG(ctx, end <= 0x7fffffffLL, "trace_ion_yield_batch: more than 2^31-1 rays");
    d.end = (unsigned)end;
    if (a.bvh->nf > max_nf) max_nf = a.bvh->nf;
  }
  if (ctx->deterministic) {
    long long total = 0;
    for (int f = 0; f < nfeat; ++f) total += 5LL * feats[f].bvh->nf;
    unsigned long long* acc = det_accumulators(ctx, 1, total);
    if (!acc) return PETCH_E_HIP;
    long long off = 0;
    for (int f = 0; f < nfeat; ++f) {
      const int nf = feats[f].bvh->nf;
      for (int k = 0; k < 5; ++k) t[f].out[k] = (float*)(acc + off + (long long)k * nf);
      t[f].fixed = 1;
      off += 5LL * nf;
    }
  }
  const IonFeature* dev =
      (const IonFeature*)petch_upload_table(ctx, t.data(), t.size() * sizeof(IonFeature));
  if (!dev) return PETCH_E_HIP;
  const unsigned gx = petch_blocks(max_nf, kBS) < 64 ? petch_blocks(max_nf, kBS) : 64;
  if (max_nf > 0 && !ctx->deterministic) {
    hipLaunchKernelGGL(k_zero_ion_batch, dim3(gx, nfeat), dim3(kBS), 0, ctx->stream, dev);
    PETCH_LAUNCH_CHECK(ctx);
  }
  if (end > 0) {
    RayQueue q;
    int rc = make_queue(ctx, (int)end, q);
    if (rc) return rc;
    IonYieldBatch pol{dev, nfeat};
    const unsigned grid =
        persistent_blocks(ctx, (const void*)k_trace_ion_yield_batch, (int)end);
    hipLaunchKernelGGL(k_trace_ion_yield_batch, dim3(grid), dim3(kBS), 0, ctx->stream, pol, q,
                       petch_hit_counter(ctx, 1));
    PETCH_LAUNCH_CHECK(ctx);
  }
  if (max_nf > 0 && ctx->deterministic) {
    hipLaunchKernelGGL(k_fix_finish_ion_batch, dim3(gx, nfeat), dim3(kBS), 0, ctx->stream, dev);
    PETCH_LAUNCH_CHECK(ctx);
  }
  return PETCH_OK;
}

extern "C" int petch_trace_legacy(petch_ctx* ctx, const petch_bvh* bvh, const petch_box* box,
                                  const petch_source* src, float sticking, int n_reemit,
                                  int specular, float cos_thr, float eta, int n, uint64_t seed,
                                  float* flux, float* angacc) {
  PETCH_ARG(ctx, ctx && bvh && box && src && flux && angacc, "trace_legacy: null pointer");
  PETCH_ARG(ctx, n >= 0 && n_reemit >= 0, "trace_legacy: bad size");
  int rc = zero(ctx, flux, bvh->nf);
  if (!rc) rc = zero(ctx, angacc, bvh->nf);
  if (rc) return rc;
  if (n == 0 || bvh->nf == 0) return PETCH_OK;
  hipLaunchKernelGGL(k_trace_legacy, dim3(petch_blocks(n, kBS)), dim3(kBS), 0, ctx->stream,
                     petch_bvh_view(bvh), make_source(src), make_box(box), sticking, n_reemit,
                     specular, cos_thr, eta, n, seed, flux, angacc);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_trace_ff(petch_ctx* ctx, const petch_bvh* bvh, const petch_box* box,
                              const float* centroids, const float* gas_normals, int nf, int n_ff,
                              uint64_t seed, int* hitface) {
  PETCH_ARG(ctx, ctx && bvh && box && centroids && gas_normals && hitface,
            "trace_ff: null pointer");
  PETCH_ARG(ctx, nf >= 0 && n_ff >= 0, "trace_ff: bad size");
  long long n = (long long)nf * n_ff;
  PETCH_ARG(ctx, n < (1LL << 31), "trace_ff: nf*n_ff too large");
  if (n == 0) return PETCH_OK;
  hipLaunchKernelGGL(k_trace_ff, dim3(petch_blocks(n, kBS)), dim3(kBS), 0, ctx->stream,
                     petch_bvh_view(bvh), make_box(box), centroids, gas_normals, nf, n_ff, seed,
                     hitface);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_set_trace_occupancy(petch_ctx* ctx, int blocks_per_cu) {
  PETCH_ARG(ctx, ctx, "set_trace_occupancy: null context");
  PETCH_ARG(ctx, blocks_per_cu >= 0, "set_trace_occupancy: blocks_per_cu < 0");
  ctx->trace_blocks_per_cu = blocks_per_cu;
  return PETCH_OK;
}

extern "C" int petch_trace_occupancy_limit(petch_ctx* ctx, int* blocks_per_cu) {
  PETCH_ARG(ctx, ctx && blocks_per_cu, "trace_occupancy_limit: null pointer");
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_trace_neutral_rr, kBS,
                                                   0) != hipSuccess)
    return petch_set_error(ctx, PETCH_E_HIP, "trace_occupancy_limit: occupancy query failed");
  *blocks_per_cu = per_cu;
  return PETCH_OK;
}

extern "C" int petch_trace_hit_counts(petch_ctx* ctx, int reset, unsigned long long* counts_host) {
  PETCH_ARG(ctx, ctx, "trace_hit_counts: null context");
  unsigned long long* c = petch_hit_counter(ctx, 0);
  if (!c) return PETCH_E_HIP;
  if (counts_host) {
    PETCH_HIP(ctx, hipMemcpyAsync(counts_host, c, 2 * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, ctx->stream));
    PETCH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (reset) PETCH_HIP(ctx, hipMemsetAsync(c, 0, 4 * sizeof(unsigned long long), ctx->stream));
  return PETCH_OK;
}
