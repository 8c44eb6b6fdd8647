// Shared helpers of the HIP library: error plumbing, device buffers, launch helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace ifea {

// Error raised inside the library and turned into a C return code at the C-ABI boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);
const char* last_error();

#define IFEA_HIP(call)                                                                                   \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess)                                                                                \
      throw ::ifea::Error(-6, std::string(#call) + " failed: " + hipGetErrorString(e_));                 \
  } while (0)

// Device allocation owned by a handle; freed on destruction.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    IFEA_HIP(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    alloc(count);
    if (count) IFEA_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  size_t bytes() const { return n * sizeof(T); }
};

// Workgroups for one item per thread. A dispatch covers fewer than 2^32 work-items: larger work needs a grid-stride
// kernel (grid_stride below), and asking for it here is an error rather than a silently shortened grid.
inline unsigned grid_for(long long work, int block) {
  long long g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g * block > 0xffffffffLL) throw Error(-3, "launch of " + std::to_string(work) + " items exceeds one dispatch");
  return (unsigned)g;
}

// Workgroups of a grid-stride kernel over `work` items: at most 2^20 workgroups.
inline unsigned grid_stride(long long work, int block) {
  long long g = (work + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : (g > (1 << 20) ? (1 << 20) : g));
}

}  // namespace ifea
