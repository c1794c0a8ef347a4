// Measurement helpers: the integer VALU rate of the device, for the roofline
// of the DP kernels (every lane runs 8 independent chains of x = max(x + y,
// z), i.e. one v_add and one v_max per chain and step, in int32 or in packed
// int16: v_pk_add_u16, v_pk_max_i16, two values per lane and instruction);
// and a plain device copy of known size, whose HBM rate and profiler byte
// counts calibrate those of the other kernels.
#include "../../include/mestrado.h"
#include "common.h"
#include "ctx.h"

namespace {

typedef short s2 __attribute__((ext_vector_type(2)));

constexpr int kChains = 8;

__global__ void __launch_bounds__(256) valu_i32_kernel(const int* in, int* out, int iters) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = in[0], z0 = in[1];
  int x[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) x[k] = t + k;
#pragma unroll 4
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) x[k] = max(x[k] + y, z0 + k);
  }
  int s = 0;
#pragma unroll
  for (int k = 0; k < kChains; ++k) s ^= x[k];
  if (s == 0x7FFFFFFF) out[t] = s;  // keeps the chains alive; never true for the inputs used
}

__global__ void __launch_bounds__(256) valu_pk16_kernel(const int* in, int* out, int iters) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const s2 y = __builtin_bit_cast(s2, in[0]);
  s2 z[kChains], x[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) {
    z[k] = __builtin_bit_cast(s2, in[1] + k);
    x[k] = __builtin_bit_cast(s2, t + k);
  }
#pragma unroll 4
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) x[k] = __builtin_elementwise_max(x[k] + y, z[k]);
  }
  int s = 0;
#pragma unroll
  for (int k = 0; k < kChains; ++k) s ^= __builtin_bit_cast(int, x[k]);
  if (s == 0x7FFFFFFF) out[t] = s;
}

// dst[i] = src[i], 16 bytes per thread and iteration, grid-stride
__global__ void __launch_bounds__(256) copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                   int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

#define CK(call)                                                                      \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return ctx->fail(MST_E_HIP, #call, hipGetErrorString(e_)); \
  } while (0)

extern "C" int mst_valu_peak(mst_ctx* ctx, int packed, double* tops) {
  if (!ctx || !tops) return MST_E_ARG;
  CK(hipSetDevice(ctx->device));
  int* buf = nullptr;
  const int groups = ctx->num_cus * 8, threads = 256, iters = 1 << 14;
  CK(hipMalloc(&buf, sizeof(int) * (2 + (size_t)groups * threads)));
  const int host[2] = {1, -5};
  hipError_t e = hipMemcpy(buf, host, sizeof host, hipMemcpyHostToDevice);
  float best = 0;
  for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {
    e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e != hipSuccess) break;
    if (packed)
      hipLaunchKernelGGL(valu_pk16_kernel, dim3(groups), dim3(threads), 0, ctx->stream, buf, buf + 2, iters);
    else
      hipLaunchKernelGGL(valu_i32_kernel, dim3(groups), dim3(threads), 0, ctx->stream, buf, buf + 2, iters);
    if ((e = hipGetLastError()) != hipSuccess) break;
    if ((e = hipEventRecord(ctx->ev1, ctx->stream)) != hipSuccess) break;
    if ((e = hipEventSynchronize(ctx->ev1)) != hipSuccess) break;
    float ms = 0;
    if ((e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1)) != hipSuccess) break;
    if (rep > 0 && (best == 0 || ms < best)) best = ms;
  }
  (void)hipFree(buf);
  if (e != hipSuccess) return ctx->fail(MST_E_HIP, "mst_valu_peak", hipGetErrorString(e));
  // two instructions per chain and step; a packed instruction does two int16 ops
  const double ops = (double)groups * threads * iters * kChains * 2 * (packed ? 2 : 1);
  *tops = ops / (best * 1e-3) / 1e12;
  return MST_OK;
}

extern "C" int mst_copy_probe(mst_ctx* ctx, int64_t bytes, int reps, double* gbps) {
  if (!ctx || !gbps || bytes < 16 || reps < 1) return MST_E_ARG;
  CK(hipSetDevice(ctx->device));
  const int64_t n = bytes / 16;
  uint4 *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, sizeof(uint4) * n));
  hipError_t e = hipMalloc(&b, sizeof(uint4) * n);
  if (e == hipSuccess) e = hipMemset(a, 1, sizeof(uint4) * n);
  float best = 0;
  const int grid = ctx->num_cus * 16;
  for (int rep = 0; rep <= reps && e == hipSuccess; ++rep) {
    if ((e = hipEventRecord(ctx->ev0, ctx->stream)) != hipSuccess) break;
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, ctx->stream, a, b, n);
    if ((e = hipGetLastError()) != hipSuccess) break;
    if ((e = hipEventRecord(ctx->ev1, ctx->stream)) != hipSuccess) break;
    if ((e = hipEventSynchronize(ctx->ev1)) != hipSuccess) break;
    float ms = 0;
    if ((e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1)) != hipSuccess) break;
    if (rep > 0 && (best == 0 || ms < best)) best = ms;  // the first run warms up
  }
  (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e != hipSuccess) return ctx->fail(MST_E_HIP, "mst_copy_probe", hipGetErrorString(e));
  *gbps = 2.0 * 16 * n / (best * 1e-3) / 1e9;  // bytes read + written
  return MST_OK;
}
