// Plain streaming copy of known size: the calibration kernel for HBM counters
// and the achievable-bandwidth reference of the roofline.
#include "ostc_kernels.h"

namespace ostc {
namespace {
__global__ __launch_bounds__(256) void k_stream_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                     int64_t n) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    dst[q] = src[q];
}
}  // namespace
}  // namespace ostc

extern "C" int ostc_stream_copy(void *dst, const void *src, long bytes, void *stream) {
  if (!dst || !src || bytes <= 0 || (bytes % 16) != 0) return OSTC_ERR_ARG;
  const int64_t n = bytes / 16;
  const unsigned blocks = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
  hipLaunchKernelGGL(ostc::k_stream_copy, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const float4 *>(src), static_cast<float4 *>(dst), n);
  return hipGetLastError() == hipSuccess ? OSTC_OK : OSTC_ERR_HIP;
}

extern "C" int ostc_device_alloc(void **ptr, long bytes) {
  if (!ptr || bytes <= 0) return OSTC_ERR_ARG;
  return hipMalloc(ptr, (size_t)bytes) == hipSuccess ? OSTC_OK : OSTC_ERR_ALLOC;
}

extern "C" int ostc_device_free(void *ptr) {
  if (ptr) (void)hipFree(ptr);
  return OSTC_OK;
}

extern "C" int ostc_memcpy_h2d(void *dst, const void *src, long bytes) {
  return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice) == hipSuccess ? OSTC_OK : OSTC_ERR_HIP;
}

extern "C" int ostc_memcpy_d2h(void *dst, const void *src, long bytes) {
  return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost) == hipSuccess ? OSTC_OK : OSTC_ERR_HIP;
}
