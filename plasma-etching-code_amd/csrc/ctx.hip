// Context lifetime, stream binding, grow-only scratch and error text.
#include "common.h"

#include <stdlib.h>
#include <string.h>

int petch_set_error(petch_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

void* petch_scratch(petch_ctx* ctx, int s, size_t bytes) {
  if (s < 0 || s >= SCR_COUNT) return NULL;
  if (bytes == 0) bytes = 256;
  if (ctx->scr_cap[s] >= bytes) return ctx->scr[s];
  // grow by 1.5x so slowly growing meshes do not reallocate every step
  size_t want = bytes + bytes / 2;
  want = (want + 255) & ~(size_t)255;
  if (ctx->scr[s]) {
    // the old buffer may still be read by queued kernels
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
      petch_set_error(ctx, PETCH_E_HIP, "scratch grow: stream sync failed");
      return NULL;
    }
    (void)hipFree(ctx->scr[s]);
    ctx->scr[s] = NULL;
    ctx->scr_cap[s] = 0;
  }
  void* p = NULL;
  if (hipMalloc(&p, want) != hipSuccess) {
    petch_set_error(ctx, PETCH_E_HIP, "scratch slot %d: hipMalloc(%zu) failed", s, want);
    return NULL;
  }
  ctx->scr[s] = p;
  ctx->scr_cap[s] = want;
  return p;
}

void* petch_pinned(petch_ctx* ctx, size_t bytes) {
  if (ctx->pinned_cap >= bytes) return ctx->pinned;
  if (ctx->pinned) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipHostFree(ctx->pinned);
    ctx->pinned = NULL;
    ctx->pinned_cap = 0;
  }
  size_t want = bytes < 4096 ? 4096 : bytes;
  void* p = NULL;
  if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
    petch_set_error(ctx, PETCH_E_HIP, "hipHostMalloc(%zu) failed", want);
    return NULL;
  }
  ctx->pinned = p;
  ctx->pinned_cap = want;
  return p;
}

void* petch_upload_table(petch_ctx* ctx, const void* host, size_t bytes) {
  void* dev = petch_scratch(ctx, SCR_TABLE, bytes);
  if (!dev) return NULL;
  const int i = ctx->ring_next;
  ctx->ring_next = (i + 1) % 8;
  if (ctx->ring_ev[i] && hipEventSynchronize(ctx->ring_ev[i]) != hipSuccess) {
    petch_set_error(ctx, PETCH_E_HIP, "upload_table: event wait failed");
    return NULL;
  }
  if (ctx->ring_cap[i] < bytes) {
    if (ctx->ring[i]) (void)hipHostFree(ctx->ring[i]);
    ctx->ring[i] = NULL;
    ctx->ring_cap[i] = 0;
    size_t want = bytes < 16384 ? 16384 : bytes;
    if (hipHostMalloc(&ctx->ring[i], want, hipHostMallocDefault) != hipSuccess) {
      ctx->ring[i] = NULL;
      petch_set_error(ctx, PETCH_E_HIP, "hipHostMalloc(%zu) failed", want);
      return NULL;
    }
    ctx->ring_cap[i] = want;
  }
  if (!ctx->ring_ev[i] &&
      hipEventCreateWithFlags(&ctx->ring_ev[i], hipEventDisableTiming) != hipSuccess) {
    ctx->ring_ev[i] = NULL;
    petch_set_error(ctx, PETCH_E_HIP, "upload_table: hipEventCreate failed");
    return NULL;
  }
  memcpy(ctx->ring[i], host, bytes);
  if (hipMemcpyAsync(dev, ctx->ring[i], bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipEventRecord(ctx->ring_ev[i], ctx->stream) != hipSuccess) {
    petch_set_error(ctx, PETCH_E_HIP, "upload_table: copy failed");
    return NULL;
  }
  return dev;
}

extern "C" int petch_ctx_create(int device, petch_ctx** out) {
  if (!out) return PETCH_E_ARG;
  *out = NULL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return PETCH_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return PETCH_E_HIP;
  petch_ctx* c = (petch_ctx*)calloc(1, sizeof(petch_ctx));
  if (!c) return PETCH_E_HIP;
  c->device = device;
  c->stream = 0;
  petch_noalloc_setup(c);
  *out = c;
  return PETCH_OK;
}

extern "C" int petch_ctx_destroy(petch_ctx* ctx) {
  if (!ctx) return PETCH_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (int s = 0; s < SCR_COUNT; ++s)
    if (ctx->scr[s]) (void)hipFree(ctx->scr[s]);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  for (int i = 0; i < 8; ++i) {
    if (ctx->ring[i]) (void)hipHostFree(ctx->ring[i]);
    if (ctx->ring_ev[i]) (void)hipEventDestroy(ctx->ring_ev[i]);
  }
  free(ctx);
  return PETCH_OK;
}

extern "C" int petch_set_stream(petch_ctx* ctx, void* stream) {
  if (!ctx) return PETCH_E_ARG;
  ctx->stream = (hipStream_t)stream;
  return PETCH_OK;
}

extern "C" int petch_set_deterministic(petch_ctx* ctx, int on) {
  if (!ctx) return PETCH_E_ARG;
  ctx->deterministic = on ? 1 : 0;
  return PETCH_OK;
}

extern "C" int petch_synchronize(petch_ctx* ctx) {
  if (!ctx) return PETCH_E_ARG;
  PETCH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PETCH_OK;
}

extern "C" const char* petch_last_error(petch_ctx* ctx) {
  return ctx ? ctx->err : "null context";
}

extern "C" int petch_version(void) { return 1; }
