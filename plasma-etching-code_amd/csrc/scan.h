// Wavefront / workgroup prefix sums (64-lane CDNA wavefronts) and a device-wide exclusive scan
// built from them: per-block reduce -> one-block scan of the block sums -> per-block downsweep.
#pragma once
#include <hip/hip_runtime.h>

namespace petch {
// internal linkage: every translation unit that includes this gets its own copies
namespace {

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Exclusive scan across a workgroup of BS threads (BS multiple of 64, <= 1024); returns the
// thread's exclusive prefix and writes the workgroup total to `total`. Uses BS/64 ints of LDS.
template <int BS>
__device__ __forceinline__ int block_excl_scan(int v, int& total) {
  __shared__ int wsum[BS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < BS / 64; ++i) {
    int s = wsum[i];
    base += (i < w) ? s : 0;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return base + inc - v;
}

template <int BS>
__device__ __forceinline__ int block_sum(int v) {
  __shared__ int wsum[BS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int s = wave_sum(v);
  if (lane == 0) wsum[w] = s;
  __syncthreads();
  int tot = 0;
#pragma unroll
  for (int i = 0; i < BS / 64; ++i) tot += wsum[i];
  __syncthreads();
  return tot;
}

// In-place exclusive scan of `nb` partial sums by one workgroup of 1024 threads; the grand total
// goes to *total. `nb` may exceed 1024 (the workgroup walks it in chunks).
__global__ void __launch_bounds__(1024) k_scan_partials(int* __restrict__ part, int nb,
                                                        int* __restrict__ total) {
  int carry = 0;
  for (int c0 = 0; c0 < nb; c0 += 1024) {
    int i = c0 + threadIdx.x;
    int v = (i < nb) ? part[i] : 0;
    int tot;
    int ex = block_excl_scan<1024>(v, tot);
    if (i < nb) part[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

// Two-array variant (used where two counts are scanned together, e.g. vertices and faces).
__global__ void __launch_bounds__(1024) k_scan_partials2(int* __restrict__ a, int* __restrict__ b,
                                                         int nb, int* __restrict__ totals) {
  int ca = 0, cb = 0;
  for (int c0 = 0; c0 < nb; c0 += 1024) {
    int i = c0 + threadIdx.x;
    int va = (i < nb) ? a[i] : 0;
    int vb = (i < nb) ? b[i] : 0;
    int ta, tb;
    int ea = block_excl_scan<1024>(va, ta);
    int eb = block_excl_scan<1024>(vb, tb);
    if (i < nb) {
      a[i] = ca + ea;
      b[i] = cb + eb;
    }
    ca += ta;
    cb += tb;
  }
  if (threadIdx.x == 0) {
    totals[0] = ca;
    totals[1] = cb;
  }
}

constexpr int kScanBS = 256;

__global__ void __launch_bounds__(kScanBS) k_scan_reduce(const int* __restrict__ in, long long n,
                                                         int* __restrict__ part) {
  long long i = (long long)blockIdx.x * kScanBS + threadIdx.x;
  int v = (i < n) ? in[i] : 0;
  int s = block_sum<kScanBS>(v);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kScanBS) k_scan_down(const int* __restrict__ in, long long n,
                                                       const int* __restrict__ part,
                                                       int* __restrict__ out) {
  long long i = (long long)blockIdx.x * kScanBS + threadIdx.x;
  int v = (i < n) ? in[i] : 0;
  int tot;
  int ex = block_excl_scan<kScanBS>(v, tot);
  if (i < n) out[i] = part[blockIdx.x] + ex;
}

}  // namespace
}  // namespace petch
