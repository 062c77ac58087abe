// Device-wide exclusive scan of int32 counts (reduce -> scan of block sums -> downsweep).
#include "common.h"
#include "scan.h"

int petch_exclusive_scan(petch_ctx* ctx, const int* in, int* out, long long n, int* total_dev) {
  if (n <= 0) {
    PETCH_HIP(ctx, hipMemsetAsync(total_dev, 0, sizeof(int), ctx->stream));
    return PETCH_OK;
  }
  long long nb = (n + petch::kScanBS - 1) / petch::kScanBS;
  int* part = (int*)petch_scratch(ctx, SCR_RED, (size_t)nb * sizeof(int));
  if (!part) return PETCH_E_HIP;
  hipLaunchKernelGGL(petch::k_scan_reduce, dim3((unsigned)nb), dim3(petch::kScanBS), 0,
                     ctx->stream, in, n, part);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(petch::k_scan_partials, dim3(1), dim3(1024), 0, ctx->stream, part, (int)nb,
                     total_dev);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(petch::k_scan_down, dim3((unsigned)nb), dim3(petch::kScanBS), 0,
                     ctx->stream, in, n, part, out);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}
