// wtg_runtime.hpp -- shared plumbing of the C++ drop-in headers: one lazily
// created device-0 context per process and RAII device buffers over the C ABI.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>
#include "wtg.h"

namespace wtg {

inline void check(int rc) {
  if (rc != WTG_OK) throw std::runtime_error(std::string("wtg: ") + wtg_last_error());
}

// Process-wide context on device 0 (WTG_DEVICE selects another), created on
// first use and destroyed at exit.  The reference's calls are made from one
// host thread; concurrent callers must use their own wtg_ctx.
inline wtg_ctx* default_context() {
  static std::once_flag once;
  static wtg_ctx* ctx = nullptr;
  std::call_once(once, [] {
    const char* d = std::getenv("WTG_DEVICE");
    check(wtg_ctx_create(d ? std::atoi(d) : 0, &ctx));
    std::atexit([] { wtg_ctx_destroy(ctx); });
  });
  return ctx;
}

// owning device buffer
class dbuf {
 public:
  dbuf() = default;
  explicit dbuf(uint64_t bytes) { alloc(bytes); }
  ~dbuf() { release(); }
  dbuf(const dbuf&) = delete;
  dbuf& operator=(const dbuf&) = delete;
  dbuf(dbuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  dbuf& operator=(dbuf&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void alloc(uint64_t bytes) {
    release();
    check(wtg_malloc(default_context(), &p_, bytes ? bytes : 1));
    n_ = bytes;
  }
  void release() {
    if (p_) wtg_free(default_context(), p_);
    p_ = nullptr;
    n_ = 0;
  }
  void upload(const void* h, uint64_t bytes) { check(wtg_memcpy_h2d(default_context(), p_, h, bytes)); }
  void download(void* h, uint64_t bytes) const { check(wtg_memcpy_d2h(default_context(), h, p_, bytes)); }
  template <typename U> U* as() const { return static_cast<U*>(p_); }
  void* get() const { return p_; }
  uint64_t bytes() const { return n_; }

 private:
  void* p_ = nullptr;
  uint64_t n_ = 0;
};

}  // namespace wtg
