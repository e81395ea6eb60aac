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

// A render entry point's work on stream s: counters (numCounters 64-bit
// words, owned by the rendered object) are zeroed, then launch() enqueues
// the kernels. Without stats requested the call only enqueues: the host
// does not wait, and the counters are left for nobody. With stats, the
// launch is timed with events, the host waits for the stream and h receives
// the counters; the return value is the time in ms.
template <class F>
float render_on_stream(hipStream_t s, DevBuf<unsigned long long> &counters, int numCounters, bool wantStats,
                       unsigned long long *h, F &&launch) {
  counters.ensure(numCounters);
  BN_HIP(hipMemsetAsync(counters.ptr, 0, numCounters * sizeof(unsigned long long), s));
  if (!wantStats) {
    launch();
    BN_HIP(hipGetLastError());
    return 0.f;
  }
  hipEvent_t e0, e1;
  BN_HIP(hipEventCreate(&e0));
  BN_HIP(hipEventCreate(&e1));
  BN_HIP(hipEventRecord(e0, s));
  launch();
  BN_HIP(hipGetLastError());
  BN_HIP(hipEventRecord(e1, s));
  BN_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  BN_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  BN_HIP(hipMemcpyAsync(h, counters.ptr, numCounters * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  BN_HIP(hipStreamSynchronize(s));
  return ms;
}

}  // namespace bnamd
