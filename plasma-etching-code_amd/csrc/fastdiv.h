// Unsigned 32-bit division by a divisor fixed for a whole launch, as a multiply-high, an add and
// two shifts (Granlund & Montgomery, "Division by invariant integers using multiplication",
// PLDI 1994, Fig. 4.1): for d >= 2, l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
//   q = (t + ((n - t) >> 1)) >> (l - 1),  t = mulhi(m, n),
// which equals floor(n / d) for every 32-bit n. d = 1 is carried as m = 0, l = 0 and returns n.
// Grid kernels use it to split a linear node index into (i, j, k) without 64-bit divisions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace petch {

struct FastDiv {
  uint32_t d, m, l;
};

__host__ __device__ inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f{d, 0u, 0u};
  if (d <= 1u) return f;
  uint32_t l = 0;
  while (l < 32u && (1ull << l) < (unsigned long long)d) ++l;
  f.l = l;
  f.m = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1ull);
  return f;
}

__device__ __forceinline__ uint32_t fastdiv(uint32_t n, const FastDiv& f) {
  if (f.l == 0u) return n;
  const uint32_t t = __umulhi(f.m, n);
  return (t + ((n - t) >> 1)) >> (f.l - 1u);
}

}  // namespace petch
