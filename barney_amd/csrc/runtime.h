// Host-side runtime helpers shared by the C-ABI translation units: error
// reporting (status code + last-error string, mirroring the throw-and-catch
// pattern of rtcore/cudaCommon/cuda-helper.h:41-60 and the catch in
// bnContextCreate, barney/api/barney.cu), device buffers and streams.
#pragma once
#include <hip/hip_runtime.h>
#include <stdexcept>
#include <string>
#include <vector>
#include <cstring>

namespace bnamd {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void set_last_error(const std::string &msg);

#define BN_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess)                                                               \
      throw ::bnamd::Error(std::string(#call) + " failed: " + hipGetErrorString(_e) +   \
                           " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")");     \
  } while (0)

#define BN_REQUIRE(cond, msg)                                 \
  do {                                                        \
    if (!(cond)) throw ::bnamd::Error(std::string(msg));      \
  } while (0)

// Runs f() and converts exceptions into the C-ABI status convention.
template <typename F>
inline int guarded(F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -1;
  } catch (...) {
    set_last_error("unknown error");
    return -1;
  }
}

// Owning device allocation.
template <typename T>
struct DevBuf {
  T *ptr = nullptr;
  size_t count = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), count(o.count) { o.ptr = nullptr; o.count = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); ptr = o.ptr; count = o.count; o.ptr = nullptr; o.count = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n) {
    release();
    if (n) BN_HIP(hipMalloc((void **)&ptr, n * sizeof(T)));
    count = n;
  }
  void ensure(size_t n) {
    if (n > count) alloc(n);
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  void upload(const T *h, size_t n, hipStream_t s = 0) {
    ensure(n);
    if (n) BN_HIP(hipMemcpyAsync(ptr, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  size_t bytes() const { return count * sizeof(T); }
};

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace bnamd
