// Radiosity neutral solve (mc_flux_3d_radiosity, threed.py:923-991).
//
//   assemble  first-hit faces of the n_ff form-factor rays per face (k_trace_ff) ->
//             D_j = escapes of j / n_ff, and the matrix Aff[i,j] = (rays j -> i) / n_ff stored as
//             a CSR of destination rows whose entries are the individual rays (duplicates are not
//             merged: the row sum over rays equals the merged sparse product). Built by a stable
//             radix sort of (destination, ray id), so every row lists its sources in ray order
//             and the float64 row sums are reproducible run to run.
//   solve     G = D + Aff diag(1 - s) G with s = clip(bare*beta, 0, 1), Jacobi from G0 = D in
//             float64 (one row per lane), n_iter sweeps (the reference's 40, threed.py:965-969)
//             or until max|dG| <= tol*max|G| (the converged solution the GMRES option of
//             threed.py:973-982 approximates to 1e-3); output clip(G, 0, 8) (threed.py:985).
//   adjoint   the implicit (fixed-point) adjoint of the solve, which the reference does not have:
//             lambda = gG + diag(q) Aff^T lambda with q = 1 - s and gG = d loss / dG (the output
//             gradient where 0 <= G <= 8), then d loss / dD = lambda and
//             d loss / d bare_j = -beta [0 <= bare_j beta <= 1] G_j (Aff^T lambda)_j.
//             Aff^T needs no second sort: row j of Aff^T is face j's own n_ff form-factor rays,
//             hitface[j n_ff .. (j+1) n_ff), so a sweep is one 64-lane wave per source face
//             that gathers lambda at its rays' hit faces (coalesced hitface reads).
#include "common.h"

int petch_radix_sort_pairs(petch_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* ktmp,
                           uint32_t* vtmp, int n, int bits);

namespace {

constexpr int kBS = 256;

__global__ void k_rad_keys(const int* __restrict__ hit, long long n, int nf,
                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  long long p = (long long)blockIdx.x * kBS + threadIdx.x;
  if (p >= n) return;
  int h = hit[p];
  keys[p] = (h >= 0 && h < nf) ? (uint32_t)h : (uint32_t)nf;  // escapes sort last
  vals[p] = (uint32_t)p;
}

__global__ void k_rad_rows(const uint32_t* __restrict__ keys, long long n, int nf,
                           int* __restrict__ row_ptr) {
  int i = blockIdx.x * kBS + threadIdx.x;
  if (i > nf) return;
  // lower bound of i in the sorted keys
  long long lo = 0, hi = n;
  while (lo < hi) {
    long long mid = (lo + hi) >> 1;
    if (keys[mid] < (uint32_t)i) lo = mid + 1; else hi = mid;
  }
  row_ptr[i] = (int)lo;
}

__global__ void k_rad_cols(const uint32_t* __restrict__ vals, const int* __restrict__ row_ptr,
                           int nf, int n_ff, int* __restrict__ col) {
  long long k = (long long)blockIdx.x * kBS + threadIdx.x;
  if (k >= row_ptr[nf]) return;
  col[k] = (int)(vals[k] / (uint32_t)n_ff);
}

__global__ void k_rad_escape(const int* __restrict__ hit, int nf, int n_ff,
                             double* __restrict__ D) {
  int j = blockIdx.x * kBS + threadIdx.x;
  if (j >= nf) return;
  int c = 0;
  for (int k = 0; k < n_ff; ++k) c += (hit[(long long)j * n_ff + k] < 0);
  D[j] = (double)c / (double)n_ff;
}

__global__ void k_rad_q(const float* __restrict__ bare, double beta, int nf,
                        double* __restrict__ q) {
  int j = blockIdx.x * kBS + threadIdx.x;
  if (j >= nf) return;
  double s = fmin(fmax((double)bare[j] * beta, 0.0), 1.0);
  q[j] = 1.0 - s;
}

__global__ void k_rad_jacobi(const int* __restrict__ row_ptr, const int* __restrict__ col,
                             const double* __restrict__ D, const double* __restrict__ q,
                             const double* __restrict__ Gin, int nf, double inv_nff,
                             double* __restrict__ Gout, unsigned long long* __restrict__ red) {
  int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= nf) return;
  double acc = 0.0;
  for (int k = row_ptr[i], e = row_ptr[i + 1]; k < e; ++k) {
    int j = col[k];
    acc += q[j] * Gin[j];
  }
  double g = D[i] + inv_nff * acc;
  Gout[i] = g;
  if (red) {
    // non-negative doubles order like their bit patterns
    atomicMax(red, (unsigned long long)__double_as_longlong(fabs(g - Gin[i])));
    atomicMax(red + 1, (unsigned long long)__double_as_longlong(fabs(g)));
  }
}

__global__ void k_rad_out(const double* __restrict__ G, int nf, float* __restrict__ out) {
  int i = blockIdx.x * kBS + threadIdx.x;
  if (i >= nf) return;
  out[i] = (float)fmin(fmax(G[i], 0.0), 8.0);
}

// Adjoint set-up: q = 1 - clip(bare beta, 0, 1); gG = gm where 0 <= G <= 8 (the clip passes the
// gradient at its bounds, as torch.clamp does); lambda0 = gG.
__global__ void k_rad_adj_init(const double* __restrict__ G, const float* __restrict__ gm,
                               const float* __restrict__ bare, double beta, int nf,
                               double* __restrict__ q, double* __restrict__ gG,
                               double* __restrict__ lam) {
  int j = blockIdx.x * kBS + threadIdx.x;
  if (j >= nf) return;
  q[j] = 1.0 - fmin(fmax((double)bare[j] * beta, 0.0), 1.0);
  double g = G[j];
  double v = (g >= 0.0 && g <= 8.0) ? (double)gm[j] : 0.0;
  gG[j] = v;
  lam[j] = v;
}

// (Aff^T lambda)_j * n_ff for the source face of this wave: the sum of lambda over the hit faces
// of its n_ff rays (escapes and out-of-range ids add nothing, as they have no CSR entry). The
// lanes add in a fixed order, so the result is reproducible run to run.
__device__ __forceinline__ double rad_gather(const int* __restrict__ hit, int j, int nf, int n_ff,
                                             const double* __restrict__ lam, int lane) {
  double acc = 0.0;
  const int* row = hit + (long long)j * n_ff;
  for (int r = lane; r < n_ff; r += 64) {
    int h = row[r];
    if (h >= 0 && h < nf) acc += lam[h];
  }
  for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
  return acc;
}

// One transposed Jacobi sweep, a wave per source face: lam_out_j = gG_j + q_j (Aff^T lam_in)_j.
__global__ void k_rad_adj_jacobi(const int* __restrict__ hit, int nf, int n_ff, double inv_nff,
                                 const double* __restrict__ q, const double* __restrict__ gG,
                                 const double* __restrict__ lin, double* __restrict__ lout,
                                 unsigned long long* __restrict__ red) {
  int j = blockIdx.x * (kBS / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (j >= nf) return;  // whole waves leave together
  double acc = rad_gather(hit, j, nf, n_ff, lin, lane);
  if (lane != 0) return;
  double l = gG[j] + q[j] * inv_nff * acc;
  lout[j] = l;
  if (red) {
    atomicMax(red, (unsigned long long)__double_as_longlong(fabs(l - lin[j])));
    atomicMax(red + 1, (unsigned long long)__double_as_longlong(fabs(l)));
  }
}

// Gradients from the converged lambda: gbare_j = -beta [0 <= bare_j beta <= 1] G_j (Aff^T lam)_j.
__global__ void k_rad_adj_out(const int* __restrict__ hit, int nf, int n_ff, double inv_nff,
                              const double* __restrict__ lam, const double* __restrict__ G,
                              const float* __restrict__ bare, double beta,
                              float* __restrict__ gbare) {
  int j = blockIdx.x * (kBS / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (j >= nf) return;
  double acc = rad_gather(hit, j, nf, n_ff, lam, lane);
  if (lane != 0) return;
  double sb = (double)bare[j] * beta;
  gbare[j] = (sb >= 0.0 && sb <= 1.0) ? (float)(-beta * G[j] * inv_nff * acc) : 0.0f;
}

// Jacobi driver shared by the solve and its adjoint. sweep(in, out, red) launches one sweep
// (red is NULL or the two-word max|dx|, max|x| reduction). From the iterate in *a (the other
// buffer in *c): n_iter sweeps, or with n_iter <= 0 chunks of 50 sweeps until the last sweep
// of a chunk has max|dx| <= tol max|x| (at most 20000 sweeps). *a holds the result on return.
template <class Sweep>
int jacobi_drive(petch_ctx* ctx, Sweep sweep, double** a, double** c, unsigned long long* red,
                 int n_iter, double tol, int* done_out) {
  int done = 0;
  if (n_iter > 0) {
    for (int it = 0; it < n_iter; ++it) {
      sweep(*a, *c, (unsigned long long*)nullptr);
      PETCH_LAUNCH_CHECK(ctx);
      double* t = *a; *a = *c; *c = t;
    }
    done = n_iter;
  } else {
    unsigned long long* h = (unsigned long long*)petch_pinned(ctx, 16);
    if (!h) return PETCH_E_HIP;
    const int chunk = 50, max_iter = 20000;
    bool conv = false;
    while (done < max_iter && !conv) {
      for (int it = 0; it < chunk; ++it) {
        bool last = it == chunk - 1;
        if (last) PETCH_HIP(ctx, hipMemsetAsync(red, 0, 16, ctx->stream));
        sweep(*a, *c, last ? red : (unsigned long long*)nullptr);
        PETCH_LAUNCH_CHECK(ctx);
        double* t = *a; *a = *c; *c = t;
      }
      done += chunk;
      PETCH_HIP(ctx, hipMemcpyAsync(h, red, 16, hipMemcpyDeviceToHost, ctx->stream));
      PETCH_HIP(ctx, hipStreamSynchronize(ctx->stream));
      double dmax, gmax;
      memcpy(&dmax, &h[0], 8);
      memcpy(&gmax, &h[1], 8);
      conv = dmax <= tol * gmax;
    }
  }
  *done_out = done;
  return PETCH_OK;
}

}  // namespace

extern "C" int petch_radiosity_assemble(petch_ctx* ctx, const int* hitface, int nf, int n_ff,
                                        int* row_ptr, int* col, double* D) {
  PETCH_ARG(ctx, ctx && hitface && row_ptr && col && D, "radiosity_assemble: null pointer");
  PETCH_ARG(ctx, nf >= 0 && n_ff > 0, "radiosity_assemble: bad size");
  long long n = (long long)nf * n_ff;
  PETCH_ARG(ctx, n < (1LL << 31), "radiosity_assemble: nf*n_ff too large");
  if (nf == 0) {
    PETCH_HIP(ctx, hipMemsetAsync(row_ptr, 0, sizeof(int), ctx->stream));
    return PETCH_OK;
  }
  uint32_t* keys = (uint32_t*)petch_scratch(ctx, SCR_RAD_A, (size_t)n * 8);
  uint32_t* tmp = (uint32_t*)petch_scratch(ctx, SCR_RAD_B, (size_t)n * 8);
  if (!keys || !tmp) return PETCH_E_HIP;
  uint32_t* vals = keys + n;
  uint32_t* ktmp = tmp;
  uint32_t* vtmp = tmp + n;
  hipLaunchKernelGGL(k_rad_keys, dim3(petch_blocks(n, kBS)), dim3(kBS), 0, ctx->stream, hitface, n,
                     nf, keys, vals);
  PETCH_LAUNCH_CHECK(ctx);
  int bits = 1;
  while ((1u << bits) <= (unsigned)nf) ++bits;
  int rc = petch_radix_sort_pairs(ctx, keys, vals, ktmp, vtmp, (int)n, bits);
  if (rc != PETCH_OK) return rc;
  hipLaunchKernelGGL(k_rad_rows, dim3(petch_blocks(nf + 1, kBS)), dim3(kBS), 0, ctx->stream, keys,
                     n, nf, row_ptr);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_rad_cols, dim3(petch_blocks(n, kBS)), dim3(kBS), 0, ctx->stream, vals,
                     row_ptr, nf, n_ff, col);
  PETCH_LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_rad_escape, dim3(petch_blocks(nf, kBS)), dim3(kBS), 0, ctx->stream, hitface,
                     nf, n_ff, D);
  PETCH_LAUNCH_CHECK(ctx);
  return PETCH_OK;
}

extern "C" int petch_radiosity_solve(petch_ctx* ctx, const int* row_ptr, const int* col,
                                     const double* D, int nf, int n_ff, const float* bare,
                                     double beta, int n_iter, double tol, double* G, float* out,
                                     int* iters_host) {
  PETCH_ARG(ctx, ctx && row_ptr && col && D && bare && G && out, "radiosity_solve: null pointer");
  PETCH_ARG(ctx, nf >= 0 && n_ff > 0, "radiosity_solve: bad size");
  if (iters_host) *iters_host = 0;
  if (nf == 0) return PETCH_OK;
  double* w = (double*)petch_scratch(ctx, SCR_RAD_C, (size_t)nf * 16 + 64);
  if (!w) return PETCH_E_HIP;
  double* q = w;
  double* G2 = w + nf;
  unsigned long long* red = (unsigned long long*)(w + 2 * (size_t)nf);
  dim3 g(petch_blocks(nf, kBS)), b(kBS);
  hipLaunchKernelGGL(k_rad_q, g, b, 0, ctx->stream, bare, beta, nf, q);
  PETCH_LAUNCH_CHECK(ctx);
  PETCH_HIP(ctx, hipMemcpyAsync(G, D, (size_t)nf * 8, hipMemcpyDeviceToDevice, ctx->stream));
  double inv = 1.0 / (double)n_ff;
  double* a = G;
  double* c = G2;
  int done = 0;
  auto sweep = [&](double* in, double* out_, unsigned long long* r) {
    hipLaunchKernelGGL(k_rad_jacobi, g, b, 0, ctx->stream, row_ptr, col, D, q, in, nf, inv, out_,
                       r);
  };
  int rc = jacobi_drive(ctx, sweep, &a, &c, red, n_iter, tol, &done);
  if (rc != PETCH_OK) return rc;
  if (a != G)
    PETCH_HIP(ctx, hipMemcpyAsync(G, a, (size_t)nf * 8, hipMemcpyDeviceToDevice, ctx->stream));
  hipLaunchKernelGGL(k_rad_out, g, b, 0, ctx->stream, G, nf, out);
  PETCH_LAUNCH_CHECK(ctx);
  if (iters_host) *iters_host = done;
  return PETCH_OK;
}

extern "C" int petch_radiosity_adjoint(petch_ctx* ctx, const int* hitface, int nf, int n_ff,
                                       const double* G, const float* bare, double beta,
                                       const float* gm, int n_iter, double tol, double* gD,
                                       float* gbare, int* iters_host) {
  PETCH_ARG(ctx, ctx && hitface && G && bare && gm && gD && gbare,
            "radiosity_adjoint: null pointer");
  PETCH_ARG(ctx, nf >= 0 && n_ff > 0, "radiosity_adjoint: bad size");
  PETCH_ARG(ctx, (long long)nf * n_ff < (1LL << 31), "radiosity_adjoint: nf*n_ff too large");
  PETCH_ARG(ctx, (const void*)gD != (const void*)G,
            "radiosity_adjoint: gD is a sweep buffer and must not alias G");
  if (iters_host) *iters_host = 0;
  if (nf == 0) return PETCH_OK;
  double* w = (double*)petch_scratch(ctx, SCR_RAD_C, (size_t)nf * 24 + 64);
  if (!w) return PETCH_E_HIP;
  double* q = w;
  double* gG = w + nf;
  double* L2 = w + 2 * (size_t)nf;
  unsigned long long* red = (unsigned long long*)(w + 3 * (size_t)nf);
  dim3 g(petch_blocks(nf, kBS)), b(kBS), gw(petch_blocks(nf, kBS / 64));
  hipLaunchKernelGGL(k_rad_adj_init, g, b, 0, ctx->stream, G, gm, bare, beta, nf, q, gG, gD);
  PETCH_LAUNCH_CHECK(ctx);
  double inv = 1.0 / (double)n_ff;
  double* a = gD;
  double* c = L2;
  int done = 0;
  auto sweep = [&](double* in, double* out, unsigned long long* r) {
    hipLaunchKernelGGL(k_rad_adj_jacobi, gw, b, 0, ctx->stream, hitface, nf, n_ff, inv, q, gG,
                       in, out, r);
  };
  int rc = jacobi_drive(ctx, sweep, &a, &c, red, n_iter, tol, &done);
  if (rc != PETCH_OK) return rc;
  if (a != gD)
    PETCH_HIP(ctx, hipMemcpyAsync(gD, a, (size_t)nf * 8, hipMemcpyDeviceToDevice, ctx->stream));
  hipLaunchKernelGGL(k_rad_adj_out, gw, b, 0, ctx->stream, hitface, nf, n_ff, inv, gD, G, bare,
                     beta, gbare);
  PETCH_LAUNCH_CHECK(ctx);
  if (iters_host) *iters_host = done;
  return PETCH_OK;
}
