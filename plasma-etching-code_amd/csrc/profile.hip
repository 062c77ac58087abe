// Cross-section profile of the level set: the gas area of every z slice, A(z) = dx^2 * sum over
// the slice's nodes of H_eps(-phi), with the smoothed Heaviside
//   H_eps(x) = 0 for x < -eps, 1 for x > eps, (1 + x/eps + sin(pi x/eps)/pi) / 2 between,
// so A is differentiable in phi. The equivalent radius of a hole is r(z) = sqrt(A(z)/pi); a
// profile loss on r(z) (sum over z of (r - r*)^2) reaches phi through the adjoint below. The
// reference has no profile extraction (its depth metrics, threed.py:1798-1869, are whole-cell
// column floors).
//
//   k_profile_area          one workgroup per slice: strided over the slice's columns, a
//                           fixed-order float64 reduction (deterministic), no atomics
//   k_profile_area_adjoint  one thread per node: gphi = -garea(z) dx^2 H_eps'(-phi)
#include "common.h"
#include <math.h>

namespace {

constexpr int kBS = 256;
constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ double heaviside(double x, double eps) {
  if (x <= -eps) return 0.0;
  if (x >= eps) return 1.0;
  return 0.5 * (1.0 + x / eps + sin(kPi * x / eps) / kPi);
}

__device__ __forceinline__ double heaviside_d(double x, double eps) {
  if (x <= -eps || x >= eps) return 0.0;
  return 0.5 * (1.0 + cos(kPi * x / eps)) / eps;
}

__global__ void k_profile_area(const float* __restrict__ phi, int nx, int ny, int nz, double dx2,
                               double eps, double* __restrict__ area) {
  __shared__ double part[kBS];
  const int k = blockIdx.x;
  const long long cols = (long long)nx * ny;
  double acc = 0.0;
  for (long long c = threadIdx.x; c < cols; c += kBS)
    acc += heaviside(-(double)phi[c * nz + k], eps);
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) area[k] = part[0] * dx2;
}

__global__ void k_profile_area_adjoint(const float* __restrict__ phi, long long N, int nz,
                                       double dx2, double eps, const double* __restrict__ garea,
                                       float* __restrict__ gphi) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int k = (int)(n % nz);
  gphi[n] = (float)(-garea[k] * dx2 * heaviside_d(-(double)phi[n], eps));
}

}  // namespace

extern "C" int petch_profile_area(petch_ctx* ctx, const float* phi, int nx, int ny, int nz,
                                  double dx, double eps, double* area) {
  PETCH_ARG(ctx, ctx && phi && area, "profile_area: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "profile_area: bad shape %d,%d,%d", nx, ny, nz);
  PETCH_ARG(ctx, dx > 0 && eps > 0, "profile_area: dx and eps must be > 0");
  hipLaunchKernelGGL(k_profile_area, dim3(nz), dim3(kBS), 0, ctx->stream, phi, nx, ny, nz,
                     dx * dx, eps, area);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_profile_area_adjoint(petch_ctx* ctx, const float* phi, int nx, int ny,
                                          int nz, double dx, double eps, const double* garea,
                                          float* gphi) {
  PETCH_ARG(ctx, ctx && phi && garea && gphi, "profile_area_adjoint: null pointer");
  PETCH_ARG(ctx, nx > 0 && ny > 0 && nz > 0, "profile_area_adjoint: bad shape %d,%d,%d", nx, ny,
            nz);
  PETCH_ARG(ctx, dx > 0 && eps > 0, "profile_area_adjoint: dx and eps must be > 0");
  PETCH_ARG(ctx, (const void*)gphi != (const void*)phi, "profile_area_adjoint: gphi aliases phi");
  const long long N = (long long)nx * ny * nz;
  hipLaunchKernelGGL(k_profile_area_adjoint, dim3(petch_blocks(N, kBS)), dim3(kBS), 0,
                     ctx->stream, phi, N, nz, dx * dx, eps, garea, gphi);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}
