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

__global__ void __launch_bounds__(kRsBS) k_rs_hist(const uint32_t* __restrict__ keys, int n,
                                                   int shift, int nblocks,
                                                   int* __restrict__ hist) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  long long base = (long long)blockIdx.x * kRsTile;
#pragma unroll
  for (int r = 0; r < kRsIPT; ++r) {
    long long i = base + r * kRsBS + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[(long long)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kRsBS) k_rs_scatter(const uint32_t* __restrict__ kin,
                                                      const uint32_t* __restrict__ vin,
                                                      uint32_t* __restrict__ kout,
                                                      uint32_t* __restrict__ vout, int n,
                                                      int shift, int nblocks,
                                                      const int* __restrict__ offs) {
  __shared__ int base_d[256];
  __shared__ int wc[kRsBS / 64][256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  base_d[tid] = offs[(long long)tid * nblocks + blockIdx.x];
  const unsigned long long lt = (lane == 0) ? 0ull : ((1ull << lane) - 1ull);
  long long base = (long long)blockIdx.x * kRsTile;
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

}  // namespace

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
