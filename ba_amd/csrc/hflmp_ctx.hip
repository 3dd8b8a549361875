// Platform context and device memory (replaces init_platform/stop_platform,
// linear_algebra/lmp/platform_management.cc:144-185, and the mem_cuda
// helpers, linear_algebra/lmp/cuda/mem_cuda.cu:42-135).
#include "hflmp_internal.hpp"

namespace {

__global__ void __launch_bounds__(kBlock) fill_zero_kernel(unsigned char *p, size_t bytes) {
  // 16-byte stores for the aligned body, bytes for the ragged ends.
  const size_t tid = (size_t)blockIdx.x * kBlock + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * kBlock;
  const uintptr_t addr = reinterpret_cast<uintptr_t>(p);
  const size_t head = ((16 - (addr & 15)) & 15) < bytes ? ((16 - (addr & 15)) & 15) : bytes;
  const size_t body = (bytes - head) / 16;
  uint4 *q = reinterpret_cast<uint4 *>(p + head);
  for (size_t i = tid; i < body; i += stride) q[i] = make_uint4(0, 0, 0, 0);
  const size_t tail0 = head + body * 16;
  for (size_t i = tid; i < head; i += stride) p[i] = 0;
  for (size_t i = tail0 + tid; i < bytes; i += stride) p[i] = 0;
}

}  // namespace

int fill_zero_on(hflmp_ctx *ctx, void *dptr, size_t bytes, hipStream_t s) {
  if (bytes == 0) return HFLMP_OK;
  HFLMP_REQUIRE(ctx, dptr, "fill_zero: null pointer");
  const unsigned g = grid_for((long long)((bytes + 15) / 16), 4);
  fill_zero_kernel<<<g, kBlock, 0, s>>>(static_cast<unsigned char *>(dptr), bytes);
  HFLMP_LAUNCH_CHECK(ctx);
  return HFLMP_OK;
}

extern "C" {

int hflmp_device_count(int *count) {
  if (!count) return HFLMP_EINVAL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return HFLMP_OK;
}

int hflmp_ctx_create(int device, hflmp_ctx **out) {
  if (!out) return HFLMP_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return HFLMP_EHIP;
  if (device < 0 || device >= n) return HFLMP_EINVAL;
  hflmp_ctx *ctx = new hflmp_ctx();
  ctx->device = device;
  auto fail = [&](int code) {
    hflmp_ctx_destroy(ctx);
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return fail(HFLMP_EHIP);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(HFLMP_EHIP);
  ctx->num_cus = prop.multiProcessorCount;
  // Default-priority stream (no priority is ever requested).
  if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess)
    return fail(HFLMP_EHIP);
  ctx->stream = ctx->own_stream;
  if (hipMalloc(&ctx->red_partials, 16 * (size_t)kMaxRedBlocks) != hipSuccess)
    return fail(HFLMP_ENOMEM);
  if (hipMalloc(reinterpret_cast<void **>(&ctx->red_counter), 64 * sizeof(unsigned int)) !=
      hipSuccess)
    return fail(HFLMP_ENOMEM);
  if (hipMalloc(reinterpret_cast<void **>(&ctx->red_result), 64 * sizeof(double)) != hipSuccess)
    return fail(HFLMP_ENOMEM);
  if (hipHostMalloc(&ctx->host_pinned, 4096, hipHostMallocDefault) != hipSuccess)
    return fail(HFLMP_ENOMEM);
  // Counters start at zero; the reductions leave them at zero.
  fill_zero_kernel<<<1, kBlock, 0, ctx->stream>>>(
      reinterpret_cast<unsigned char *>(ctx->red_counter), 64 * sizeof(unsigned int));
  if (hipGetLastError() != hipSuccess) return fail(HFLMP_EHIP);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(HFLMP_EHIP);
  *out = ctx;
  return HFLMP_OK;
}

int hflmp_ctx_destroy(hflmp_ctx *ctx) {
  if (!ctx) return HFLMP_EINVAL;
  (void)hipSetDevice(ctx->device);
  if (ctx->own_stream) (void)hipStreamSynchronize(ctx->own_stream);
  if (ctx->stream && ctx->stream != ctx->own_stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->red_partials) (void)hipFree(ctx->red_partials);
  if (ctx->red_counter) (void)hipFree(ctx->red_counter);
  if (ctx->red_result) (void)hipFree(ctx->red_result);
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return HFLMP_OK;
}

int hflmp_ctx_sync(hflmp_ctx *ctx) {
  if (!ctx) return HFLMP_EINVAL;
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return HFLMP_OK;
}

const char *hflmp_ctx_last_error(const hflmp_ctx *ctx) {
  return ctx ? ctx->last_error : "null context";
}

int hflmp_ctx_get_stream(hflmp_ctx *ctx, void **stream) {
  if (!ctx || !stream) return HFLMP_EINVAL;
  *stream = ctx->stream;
  return HFLMP_OK;
}

int hflmp_ctx_set_stream(hflmp_ctx *ctx, void *stream) {
  if (!ctx) return HFLMP_EINVAL;
  ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
  return HFLMP_OK;
}

int hflmp_ctx_device(const hflmp_ctx *ctx, int *device) {
  if (!ctx || !device) return HFLMP_EINVAL;
  *device = ctx->device;
  return HFLMP_OK;
}

int hflmp_event_create(hflmp_ctx *ctx, void **event) {
  if (!ctx || !event) return HFLMP_EINVAL;
  hipEvent_t e;
  HFLMP_HIP_CHECK(ctx, hipEventCreate(&e));
  *event = e;
  return HFLMP_OK;
}

int hflmp_event_record(hflmp_ctx *ctx, void *event) {
  if (!ctx || !event) return HFLMP_EINVAL;
  HFLMP_HIP_CHECK(ctx, hipEventRecord(static_cast<hipEvent_t>(event), ctx->stream));
  return HFLMP_OK;
}

int hflmp_event_elapsed_ms(hflmp_ctx *ctx, void *start, void *stop, float *ms) {
  if (!ctx || !start || !stop || !ms) return HFLMP_EINVAL;
  HFLMP_HIP_CHECK(ctx, hipEventSynchronize(static_cast<hipEvent_t>(stop)));
  HFLMP_HIP_CHECK(ctx, hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                           static_cast<hipEvent_t>(stop)));
  return HFLMP_OK;
}

int hflmp_event_destroy(hflmp_ctx *ctx, void *event) {
  if (!ctx || !event) return HFLMP_EINVAL;
  HFLMP_HIP_CHECK(ctx, hipEventDestroy(static_cast<hipEvent_t>(event)));
  return HFLMP_OK;
}

int hflmp_malloc(hflmp_ctx *ctx, size_t bytes, void **dptr) {
  if (!ctx || !dptr) return HFLMP_EINVAL;
  *dptr = nullptr;
  if (bytes == 0) return HFLMP_OK;
  HFLMP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipError_t e = hipMalloc(dptr, bytes);
  if (e != hipSuccess) {
    *dptr = nullptr;
    return set_error(ctx, HFLMP_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  return HFLMP_OK;
}

int hflmp_free(hflmp_ctx *ctx, void *dptr) {
  if (!ctx) return HFLMP_EINVAL;
  if (!dptr) return HFLMP_OK;
  // Work queued on the stream may still read the buffer.
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  HFLMP_HIP_CHECK(ctx, hipFree(dptr));
  return HFLMP_OK;
}

int hflmp_memcpy_h2d(hflmp_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return HFLMP_EINVAL;
  if (bytes == 0) return HFLMP_OK;
  HFLMP_REQUIRE(ctx, dst && src, "memcpy_h2d: null pointer");
  // Source is pageable host memory the caller may reuse on return.
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return HFLMP_OK;
}

int hflmp_memcpy_d2h(hflmp_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return HFLMP_EINVAL;
  if (bytes == 0) return HFLMP_OK;
  HFLMP_REQUIRE(ctx, dst && src, "memcpy_d2h: null pointer");
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HFLMP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return HFLMP_OK;
}

int hflmp_memcpy_d2d(hflmp_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return HFLMP_EINVAL;
  if (bytes == 0) return HFLMP_OK;
  HFLMP_REQUIRE(ctx, dst && src, "memcpy_d2d: null pointer");
  HFLMP_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return HFLMP_OK;
}

int hflmp_fill_zero(hflmp_ctx *ctx, void *dptr, size_t bytes) {
  if (!ctx) return HFLMP_EINVAL;
  return fill_zero_on(ctx, dptr, bytes, ctx->stream);
}

}  // extern "C"
