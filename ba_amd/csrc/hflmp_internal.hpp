// Internal definitions shared by the HIP translation units of libhflmp_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <type_traits>

#include "hflmp_hip.h"

// Threads per workgroup for every kernel of the library: four wave64s.
constexpr int kBlock = 256;
constexpr int kWave = 64;
// gfx950: 256 CUs in 8 XCDs, workgroups dispatched round-robin over XCDs.
constexpr int kXcds = 8;
// Upper bound on reduction partials (one per workgroup of a reduction).
constexpr int kMaxRedBlocks = 2048;

struct hflmp_ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  // Reduction scratch: kMaxRedBlocks partials (16 B each), a ticket
  // counter per concurrent reduction slot, result slots.
  void *red_partials = nullptr;
  unsigned int *red_counter = nullptr;
  double *red_result = nullptr;   // 64 device result slots
  void *host_pinned = nullptr;    // 4 KiB pinned host staging
  char last_error[512] = {0};
};

// Record an error message in the context and return the code.
inline int set_error(hflmp_ctx *ctx, int code, const char *fmt, ...) {
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->last_error, sizeof(ctx->last_error), fmt, ap);
    va_end(ap);
  }
  return code;
}

#define HFLMP_HIP_CHECK(ctx, expr)                                            \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return set_error((ctx), HFLMP_EHIP, "%s: %s (%s:%d)", #expr,            \
                       hipGetErrorString(e_), __FILE__, __LINE__);             \
  } while (0)

#define HFLMP_LAUNCH_CHECK(ctx)                                               \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess)                                                      \
      return set_error((ctx), HFLMP_EHIP, "kernel launch: %s (%s:%d)",        \
                       hipGetErrorString(e_), __FILE__, __LINE__);             \
  } while (0)

#define HFLMP_REQUIRE(ctx, cond, ...)                                         \
  do {                                                                         \
    if (!(cond)) return set_error((ctx), HFLMP_EINVAL, __VA_ARGS__);          \
  } while (0)

inline int check_ctx(hflmp_ctx *ctx) { return ctx ? HFLMP_OK : HFLMP_EINVAL; }

// hflmp_fill_zero on a given stream (errors recorded in ctx).
int fill_zero_on(hflmp_ctx *ctx, void *dptr, size_t bytes, hipStream_t s);

// Grid for an elementwise kernel that handles `per_thread` items/thread.
inline unsigned grid_for(long long n, int per_thread) {
  long long g = (n + (long long)kBlock * per_thread - 1) / ((long long)kBlock * per_thread);
  if (g < 1) g = 1;
  if (g > 65535LL * 64) g = 65535LL * 64;
  return (unsigned)g;
}

// Number of workgroups for a grid-stride reduction over n items.
inline int red_blocks(hflmp_ctx *ctx, long long n) {
  long long g = (n + 4LL * kBlock - 1) / (4LL * kBlock);
  long long cap = ctx->num_cus > 0 ? 8LL * ctx->num_cus : 1024;
  if (cap > kMaxRedBlocks) cap = kMaxRedBlocks;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

// ---- device helpers --------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sum over the workgroup; the result is valid in thread 0.  `lds` must hold
// kBlock / kWave values.  Fixed order, so the result is deterministic.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  __syncthreads();  // a previous block_sum may still be reading lds
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) s += lds[w];
  }
  return s;
}

// Grid-wide sums are two launches: every workgroup of the producing kernel
// stores its partial with a plain store, and a one-workgroup kernel sums
// the partials in index order.  The kernel boundary makes the partials
// visible; no device-scope fence (an L2 write-back on gfx950's per-XCD L2s)
// is issued per workgroup, and the result does not depend on timing.

// Threads of the one-workgroup kernels that sum partials.
constexpr int kFinalBlock = 1024;

// Sum of `count` strided partials src[k*stride + off] over one workgroup of
// kFinalBlock threads; valid in thread 0.  Eight independent accumulators
// keep eight loads in flight per thread (one round trip for up to 8192
// partials); the order is fixed, so the result is deterministic.  `lds` holds kFinalBlock / kWave values.
template <typename T>
__device__ __forceinline__ T sum_partials(const T *src, int count, int stride, int off, T *lds) {
  constexpr int F = 8;
  T a[F];
#pragma unroll
  for (int f = 0; f < F; ++f) a[f] = 0;
  int k = threadIdx.x;
  for (; k + (F - 1) * kFinalBlock < count; k += F * kFinalBlock) {
    T v[F];
#pragma unroll
    for (int f = 0; f < F; ++f) v[f] = src[(size_t)(k + f * kFinalBlock) * stride + off];
#pragma unroll
    for (int f = 0; f < F; ++f) a[f] += v[f];
  }
  // the remainder: up to F - 1 loads, all issued before any add
  {
    T v[F];
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const int kk = k + f * kFinalBlock;
      v[f] = kk < count ? src[(size_t)kk * stride + off] : T(0);
    }
#pragma unroll
    for (int f = 0; f < F; ++f) a[f] += v[f];
  }
  T v = wave_sum(((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7])));
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  __syncthreads();  // a previous call may still be reading lds
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kFinalBlock / kWave; ++w) s += lds[w];
  }
  return s;
}

// XCD-aware tile index: workgroup b runs on XCD b % 8; give each XCD a
// contiguous range of tiles so neighbouring tiles share that XCD's L2.
__device__ __forceinline__ long long xcd_tile(long long b, long long nblocks) {
  const long long per = (nblocks + kXcds - 1) / kXcds;
  return (b % kXcds) * per + b / kXcds;
}

template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
  return __builtin_nontemporal_load(p);
}
