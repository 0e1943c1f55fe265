// Host-side helpers shared by the C-ABI implementation files.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace rtpg {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);
#define RTPG_HIP(x) ::rtpg::hip_check((x), #x)

void set_last_error(const std::string& s);

// Owning device buffer of T.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // Allocates exactly `count` elements (contents undefined).
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    RTPG_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    n = count;
  }
  // Grows to at least `count` elements, discarding the contents.
  void reserve(size_t count) {
    if (count > n) alloc(count);
  }
  void upload(const T* src, size_t count, hipStream_t s) {
    if (count > n) alloc(count);
    if (count) RTPG_HIP(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  void download(T* dst, size_t count, hipStream_t s) const {
    if (count) RTPG_HIP(hipMemcpyAsync(dst, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    RTPG_HIP(hipStreamSynchronize(s));
  }
  std::vector<T> to_vector(size_t count, hipStream_t s) const {
    std::vector<T> v(count);
    download(v.data(), count, s);
    return v;
  }
  void zero(size_t count, hipStream_t s) {
    if (count) RTPG_HIP(hipMemsetAsync(p, 0, count * sizeof(T), s));
  }
};

}  // namespace rtpg
