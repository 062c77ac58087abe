// Shared host/device plumbing for libpetch_hip: the context object, grow-only scratch,
// error reporting and launch helpers. Every kernel in the library is launched on ctx->stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>
#include <string.h>

#include "../../include/petch_hip.h"

// Scratch slots owned by a context (grow-only device buffers, reused step to step).
enum PetchScratch {
  SCR_LS_A = 0,      // level set ping-pong / Eikonal U
  SCR_LS_B,          // second ping-pong buffer
  SCR_LS_FROZEN,     // Eikonal frozen flags (u8)
  SCR_MC_CELL,       // marching cubes: per-node triangle counts (u8) and vertex-slot flags
  SCR_MC_BLOCK,      // marching cubes: per-block partial sums
  SCR_ADJ,           // face adjacency: vertex->face incidence
  SCR_ADJ2,
  SCR_RED,           // small reductions / counters
  SCR_TRACE,         // ray-queue counters
  SCR_SORT_A,        // radix sort buffers
  SCR_SORT_B,
  SCR_SORT_C,
  SCR_SORT_D,
  SCR_SORT_HIST,
  SCR_RAD_A,         // radiosity CSR build
  SCR_RAD_B,
  SCR_RAD_C,
  SCR_SMOOTH,        // flux smoothing ping-pong
  SCR_TABLE,         // per-launch descriptor tables (batched traces)
  SCR_DET,           // int64 fixed-point accumulators of the deterministic mode
  SCR_EIK_STATES,    // reinit adjoint: every sweep's Eikonal state
  SCR_EIK_W,         // reinit adjoint: per-node Godunov weights of one sweep
  SCR_EIK_G,         // reinit adjoint: adjoint ping-pong
  SCR_LS_BRICK,      // batched level-set stages: per-brick activity flags and list counters
  SCR_LS_LIST,       // batched level-set stages: lists of the bricks a launch processes
  SCR_SORT_MANY,     // batched radix sort: per-array digit tables
  SCR_REDEP,         // redeposition: float64 emission CDF and its workgroup partials
  SCR_COUNT
};

struct petch_ctx {
  int device;
  hipStream_t stream;
  char err[1024];
  void* scr[SCR_COUNT];
  size_t scr_cap[SCR_COUNT];
  void* pinned;            // small pinned host staging area for counts / scalars
  size_t pinned_cap;
  // ring of pinned staging buffers for descriptor tables: slot i is refilled only after the
  // copy that read it last (event i) has run, so the host never overwrites data in flight
  void* ring[8];
  size_t ring_cap[8];
  hipEvent_t ring_ev[8];
  int ring_next;
  // this device's copy of the fixed workspace of petch_reinit_fsm_noalloc (its device address,
  // from hipGetSymbolAddress at context creation), or NULL with the error that prevented it
  void* noalloc_scratch;
  hipError_t noalloc_err;
  int deterministic;       // traces accumulate in int64 fixed point (petch_set_deterministic)
  int trace_blocks_per_cu; // persistent trace grids: workgroups per CU (0: the occupancy limit)
};

struct petch_bvh {
  int device;
  int nf, nv;              // faces / vertices of the mesh it was built over
  int cap_f;               // capacity (faces) of the device buffers below
  void* nodes;             // (nf-1) internal nodes, 64 B each: child AABBs + child refs
  void* wide;              // 4-wide nodes, 128 B each, at the indices of the binary nodes they replace
  void* tris;              // nf triangles in leaf order, 48 B each (v0, v1, v2 as float4)
  void* work;              // build scratch: morton keys/values, parents, flags
  void* aux;               // clustering scratch and tree of the large-mesh build
  size_t aux_cap;
  size_t work_cap;
  int quality;             // build quality: -1 by mesh size, 0 Morton + fixed collapse,
                           // 1 Morton + area collapse, 2 clustered (PLOC) + area collapse
  float root_lo[3], root_hi[3];
};

int petch_set_error(petch_ctx* ctx, int code, const char* fmt, ...);

// Records in ctx the device address of the current device's fixed no-allocation workspace
// (levelset.hip); called by petch_ctx_create with the context's device current.
void petch_noalloc_setup(petch_ctx* ctx);

// Copies `bytes` of host data into the SCR_TABLE device buffer on ctx->stream (stream-ordered:
// kernels launched afterwards on the stream see it; a later upload waits for them in stream
// order) and returns the device address, or NULL on failure.
void* petch_upload_table(petch_ctx* ctx, const void* host, size_t bytes);

// Returns a device buffer of at least `bytes` for slot `s`, growing it if needed (NULL on failure
// with the error recorded in ctx).
void* petch_scratch(petch_ctx* ctx, int s, size_t bytes);
// Pinned host staging of at least `bytes`.
void* petch_pinned(petch_ctx* ctx, size_t bytes);

// Exclusive scan of n int32 counts into out (distinct from in); the total goes to *total_dev.
// Uses the SCR_RED scratch slot.
int petch_exclusive_scan(petch_ctx* ctx, const int* in, int* out, long long n, int* total_dev);

// One array of a batched radix sort (petch_radix_sort_pairs_many): keys/vals sorted by their 32
// key bits with ktmp/vtmp as n-element scratch; hist and offs hold 256 x nblocks ints each;
// tblk0 is the array's first workgroup in the batched launches.
struct petch_sort_item {
  uint32_t *keys, *vals, *ktmp, *vtmp;
  int *hist, *offs;
  int n, nblocks;
  long long tblk0;
};
int petch_radix_sort_pairs_many(petch_ctx* ctx, const petch_sort_item* items_dev, int n_items,
                                long long tiles);
int petch_sort_tiles(int n);   // workgroups (tiles) of one array in the sort launches

#define PETCH_HIP(ctx, call)                                                              \
  do {                                                                                    \
    hipError_t _e = (call);                                                               \
    if (_e != hipSuccess)                                                                 \
      return petch_set_error((ctx), PETCH_E_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call, \
                             hipGetErrorString(_e));                                      \
  } while (0)

#define PETCH_LAUNCH_CHECK(ctx)                                                           \
  do {                                                                                    \
    hipError_t _e = hipGetLastError();                                                    \
    if (_e != hipSuccess)                                                                 \
      return petch_set_error((ctx), PETCH_E_HIP, "%s:%d kernel launch: %s", __FILE__,      \
                             __LINE__, hipGetErrorString(_e));                            \
  } while (0)

#define PETCH_ARG(ctx, cond, ...)                                                         \
  do {                                                                                    \
    if (!(cond)) return petch_set_error((ctx), PETCH_E_ARG, __VA_ARGS__);                 \
  } while (0)

static inline unsigned petch_blocks(long long n, int bs) {
  return (unsigned)((n + bs - 1) / bs);
}
