// Device/memory entry points of the C ABI (the rtc::Device memory surface of
// rtcore/cudaCommon/Device.h:34-64) and the last-error slot.
#include "runtime.h"
#include "../../include/bnamd.h"
#include <mutex>

namespace bnamd {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace bnamd

using namespace bnamd;

extern "C" const char *bnamd_last_error(void) { return g_last_error.c_str(); }

extern "C" int bnamd_init(int device) {
  return guarded([&] {
    int n = 0;
    BN_HIP(hipGetDeviceCount(&n));
    BN_REQUIRE(n > 0, "bnamd_init: no HIP device available");
    BN_REQUIRE(device >= 0 && device < n, "bnamd_init: device index out of range");
    BN_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    BN_HIP(hipGetDeviceProperties(&prop, device));
    BN_REQUIRE(std::string(prop.gcnArchName).rfind("gfx950", 0) == 0,
               std::string("bnamd_init: built for gfx950, device is ") + prop.gcnArchName);
  });
}

extern "C" int bnamd_device_malloc(void **dptr, size_t bytes) {
  return guarded([&] {
    BN_REQUIRE(dptr, "bnamd_device_malloc: NULL out pointer");
    *dptr = nullptr;
    if (bytes) BN_HIP(hipMalloc(dptr, bytes));
  });
}
extern "C" int bnamd_device_free(void *dptr) {
  return guarded([&] {
    if (dptr) BN_HIP(hipFree(dptr));
  });
}
extern "C" int bnamd_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  return guarded([&] {
    if (bytes) BN_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  });
}
extern "C" int bnamd_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  return guarded([&] {
    if (bytes) BN_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  });
}
extern "C" int bnamd_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  return guarded([&] {
    if (bytes) BN_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  });
}
extern "C" int bnamd_memset(void *dst, int value, size_t bytes) {
  return guarded([&] {
    if (bytes) BN_HIP(hipMemset(dst, value, bytes));
  });
}
extern "C" int bnamd_synchronize(void) {
  return guarded([&] { BN_HIP(hipDeviceSynchronize()); });
}
extern "C" int bnamd_memcpy_async(void *dst, const void *src, size_t bytes, void *stream) {
  return guarded([&] {
    if (bytes) BN_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, (hipStream_t)stream));
  });
}
extern "C" int bnamd_memset_async(void *dst, int value, size_t bytes, void *stream) {
  return guarded([&] {
    if (bytes) BN_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
  });
}
extern "C" int bnamd_host_malloc(void **ptr, size_t bytes) {
  return guarded([&] {
    BN_REQUIRE(ptr, "bnamd_host_malloc: NULL out pointer");
    *ptr = nullptr;
    if (bytes) BN_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
  });
}
extern "C" int bnamd_host_free(void *ptr) {
  return guarded([&] {
    if (ptr) BN_HIP(hipHostFree(ptr));
  });
}
extern "C" int bnamd_get_device(int *device) {
  return guarded([&] {
    BN_REQUIRE(device, "bnamd_get_device: NULL out pointer");
    BN_HIP(hipGetDevice(device));
  });
}
extern "C" int bnamd_set_device(int device) {
  return guarded([&] { BN_HIP(hipSetDevice(device)); });
}

namespace {
// Plain streaming copy (one float4 per thread per iteration); used to
// calibrate HBM byte counters against a kernel of known traffic.
__global__ void k_copy(const float4 *__restrict src, float4 *__restrict dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

extern "C" int bnamd_copy_kernel(void *dst, const void *src, size_t bytes, void *stream) {
  return guarded([&] {
    BN_REQUIRE(bytes % 16 == 0, "bnamd_copy_kernel: size must be a multiple of 16");
    BN_REQUIRE(dst && src, "bnamd_copy_kernel: NULL pointer");
    const size_t n = bytes / 16;
    hipStream_t s = (hipStream_t)stream;
    k_copy<<<2048, 256, 0, s>>>((const float4 *)src, (float4 *)dst, n);
    BN_HIP(hipGetLastError());
    if (!stream) BN_HIP(hipStreamSynchronize(s));
  });
}
