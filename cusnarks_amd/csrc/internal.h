// Library-internal declarations shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <mutex>
#include "../../include/csk.h"

namespace csk {

// thread-local last error message, returned by csk_last_error()
void set_error(const std::string &msg);

#define CSK_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::csk::set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
      return CSK_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

#define CSK_CHECK_LAUNCH()                                                         \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess) {                                                        \
      ::csk::set_error(std::string("kernel launch: ") + hipGetErrorString(e_));    \
      return CSK_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

// A growable device workspace owned by a context.  Callers carve it up with
// Carver; the memory is reused across calls on the same context (and stream
// order guarantees no overlap between successive calls on one stream).
struct Workspace {
  void *ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t need);
  void release();
};

struct Carver {
  char *base;
  size_t off = 0;
  explicit Carver(void *b) : base((char *)b) {}
  template <class T> T *take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T *p = (T *)(base + off);
    off += count * sizeof(T);
    return p;
  }
  static size_t round(size_t b) { return (b + 255) & ~(size_t)255; }
};

}  // namespace csk

struct csk_ctx {
  int device = 0;
  csk::Workspace ws;
  csk::Workspace ws2;  // second workspace for nested users (prover stages)
  std::mutex mu;       // serialises host-side use of the workspace
  uint32_t *g1_table = nullptr;  // fixed-base tables (k*2^(8w)*G), lazily built
  uint32_t *g2_table = nullptr;
};
