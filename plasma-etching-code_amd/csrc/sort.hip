// Stable LSD radix sort of (uint32 key, uint32 value) pairs, 8-bit digits.
//
// Per pass: (1) each workgroup histograms its 2048-key tile in LDS and writes the 256 counts
// digit-major (hist[d * nblocks + block]); (2) a device-wide exclusive scan of that table gives
// every (digit, block) its output offset; (3) each workgroup re-reads its tile in order, ranks keys
// within the tile stably (wavefront match of equal digits by 8 ballots, per-wave digit counts in
// LDS, waves combined in order) and scatters keys and values.
#include "common.h"
#include "scan.h"

namespace {

constexpr int kRsBS = 256;
constexpr int kRsIPT = 8;
constexpr int kRsTile = kRsBS * kRsIPT;

// Histogram of tile `tb` of one array (hist is digit-major over the array's nblocks tiles).
__device__ __forceinline__ void rs_hist_tile(const uint32_t* __restrict__ keys, int n, int shift,
                                             int nblocks, int* __restrict__ hist, int tb) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  long long base = (long long)tb * kRsTile;
#pragma unroll
  for (int r = 0; r < kRsIPT; ++r) {
    long long i = base + r * kRsBS + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[(long long)threadIdx.x * nblocks + tb] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kRsBS) k_rs_hist(const uint32_t* __restrict__ keys, int n,
                                                   int shift, int nblocks,
                                                   int* __restrict__ hist) {
  rs_hist_tile(keys, n, shift, nblocks, hist, blockIdx.x);
}

// Stable scatter of tile `tb` of one array to its digit offsets.
__device__ __forceinline__ void rs_scatter_tile(const uint32_t* __restrict__ kin,
                                                const uint32_t* __restrict__ vin,
                                                uint32_t* __restrict__ kout,
                                                uint32_t* __restrict__ vout, int n, int shift,
                                                int nblocks, const int* __restrict__ offs,
                                                int tb) {
  __shared__ int base_d[256];
  __shared__ int wc[kRsBS / 64][256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  base_d[tid] = offs[(long long)tid * nblocks + tb];
  const unsigned long long lt = (lane == 0) ? 0ull : ((1ull << lane) - 1ull);
  long long base = (long long)tb * kRsTile;
  for (int r = 0; r < kRsIPT; ++r) {
    long long i = base + r * kRsBS + tid;
    bool valid = i < n;
    uint32_t key = valid ? kin[i] : 0u;
    uint32_t val = valid ? vin[i] : 0u;
    int d = (int)((key >> shift) & 255u);
#pragma unroll
    for (int q = 0; q < kRsBS / 64; ++q) wc[q][tid] = 0;
    __syncthreads();
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      bool bit = (d >> b) & 1;
      unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    int rank = __popcll(peers & lt);
    if (valid && rank == 0) wc[w][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int pos = base_d[d] + rank;
      for (int q = 0; q < w; ++q) pos += wc[q][d];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int q = 0; q < kRsBS / 64; ++q) add += wc[q][tid];
    base_d[tid] += add;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kRsBS) k_rs_scatter(const uint32_t* __restrict__ kin,
                                                      const uint32_t* __restrict__ vin,
                                                      uint32_t* __restrict__ kout,
                                                      uint32_t* __restrict__ vout, int n,
                                                      int shift, int nblocks,
                                                      const int* __restrict__ offs) {
  rs_scatter_tile(kin, vin, kout, vout, n, shift, nblocks, offs, blockIdx.x);
}

// Many arrays sorted side by side (the BVH builds of a batched step): array a owns the
// workgroups [tblk0[a], tblk0[a+1]) of the histogram and scatter launches, and one workgroup
// per array scans its own digit-major table.
__device__ __forceinline__ int sort_item_of(const petch_sort_item* __restrict__ it, int n) {
  int lo = 0, hi = n - 1;
  const long long b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (it[mid].tblk0 <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(kRsBS) k_rs_hist_many(const petch_sort_item* __restrict__ items,
                                                        int n, int pass) {
  const petch_sort_item& it = items[sort_item_of(items, n)];
  const uint32_t* k = (pass & 1) ? it.ktmp : it.keys;
  rs_hist_tile(k, it.n, 8 * pass, it.nblocks, it.hist, (int)(blockIdx.x - it.tblk0));
}

__global__ void __launch_bounds__(kRsBS) k_rs_scatter_many(
    const petch_sort_item* __restrict__ items, int n, int pass) {
  const petch_sort_item& it = items[sort_item_of(items, n)];
  const bool odd = pass & 1;
  rs_scatter_tile(odd ? it.ktmp : it.keys, odd ? it.vtmp : it.vals, odd ? it.keys : it.ktmp,
                  odd ? it.vals : it.vtmp, it.n, 8 * pass, it.nblocks, it.offs,
                  (int)(blockIdx.x - it.tblk0));
}

// One workgroup per array: exclusive scan of its 256 x nblocks digit-major table.
__global__ void __launch_bounds__(kRsBS) k_rs_scan_many(const petch_sort_item* __restrict__ items) {
  const petch_sort_item& it = items[blockIdx.x];
  const long long m = 256LL * it.nblocks;
  int carry = 0;
  for (long long c = 0; c < m; c += kRsBS) {
    const long long i = c + threadIdx.x;
    const int v = i < m ? it.hist[i] : 0;
    int tot;
    const int ex = petch::block_excl_scan<kRsBS>(v, tot);
    if (i < m) it.offs[i] = carry + ex;
    carry += tot;
  }
}

}  // namespace

// Sorts n_items arrays of (key, value) pairs by their low 32 key bits (four 8-bit passes; the
// result ends in keys/vals), one launch per stage for all arrays. items (device table, built by
// the caller with petch_sort_plan) give each array's buffers, its tile count and first tile.
int petch_radix_sort_pairs_many(petch_ctx* ctx, const petch_sort_item* items_dev, int n_items,
                                long long tiles) {
  if (n_items <= 0 || tiles <= 0) return PETCH_OK;
  for (int pass = 0; pass < 4; ++pass) {
    hipLaunchKernelGGL(k_rs_hist_many, dim3((unsigned)tiles), dim3(kRsBS), 0, ctx->stream,
                       items_dev, n_items, pass);
    PETCH_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_rs_scan_many, dim3((unsigned)n_items), dim3(kRsBS), 0, ctx->stream,
                       items_dev);
    PETCH_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_rs_scatter_many, dim3((unsigned)tiles), dim3(kRsBS), 0, ctx->stream,
                       items_dev, n_items, pass);
    PETCH_LAUNCH_CHECK(ctx);
  }
  return PETCH_OK;
}

int petch_sort_tiles(int n) { return (n + kRsTile - 1) / kRsTile; }

// Sorts n pairs by the key bits [0, bits) (bits rounded up to a multiple of 8). Result in
// keys/vals; ktmp/vtmp are n-element scratch. Uses SCR_SORT_HIST and SCR_RED.
int petch_radix_sort_pairs(petch_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* ktmp,
                           uint32_t* vtmp, int n, int bits) {
  if (n <= 1) return PETCH_OK;
  int nblocks = (n + kRsTile - 1) / kRsTile;
  int* hist = (int*)petch_scratch(ctx, SCR_SORT_HIST, ((size_t)256 * nblocks * 2 + 1) * sizeof(int));
  if (!hist) return PETCH_E_HIP;
  int* offs = hist + 256 * nblocks;
  int* total = offs + 256 * nblocks;
  uint32_t *ka = keys, *va = vals, *kb = ktmp, *vb = vtmp;
  int passes = (bits + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    int shift = 8 * p;
    hipLaunchKernelGGL(k_rs_hist, dim3(nblocks), dim3(kRsBS), 0, ctx->stream, ka, n, shift,
                       nblocks, hist);
    PETCH_LAUNCH_CHECK(ctx);
    int rc = petch_exclusive_scan(ctx, hist, offs, 256LL * nblocks, total);
    if (rc != PETCH_OK) return rc;
    hipLaunchKernelGGL(k_rs_scatter, dim3(nblocks), dim3(kRsBS), 0, ctx->stream, ka, va, kb, vb, n,
                       shift, nblocks, offs);
    PETCH_LAUNCH_CHECK(ctx);
    uint32_t* t = ka; ka = kb; kb = t;
    t = va; va = vb; vb = t;
  }
  if (ka != keys) {
    PETCH_HIP(ctx, hipMemcpyAsync(keys, ka, (size_t)n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    PETCH_HIP(ctx, hipMemcpyAsync(vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  }
  return PETCH_OK;
}
