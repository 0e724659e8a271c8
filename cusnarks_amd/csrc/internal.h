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
  uint64_t gen = 0;  // bumped on every reallocation
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

// Launch plan of a group of J multi-exponentiations sharing one window width
// (a single call is the group J = 1).  Member j owns the bucket keys
// [j*K, (j+1)*K); every stage runs once over the whole group.
struct MsmPlan {
  uint32_t c;            // window bits
  uint32_t W;            // windows: floor(254 / c) + 1 signed-digit windows cover k < r < 2^254
  uint32_t B;            // buckets per window, 2^(c-1) (signed digits)
  uint32_t K;            // W * B bucket keys per member
  uint32_t J;            // members
  uint32_t hb;           // bucket bits the sort's first pass bins by (the second takes the other c-1-hb)
  uint32_t st;           // low bucket bits the second pass sorts the top window by (<= c-1-hb)
  uint32_t L;            // sorted entries per accumulation thread
  uint32_t Lr;           // buckets per reduction thread
  uint32_t S;            // reduction segments per window, B / Lr
  uint64_t n;            // points over all members
  uint64_t max_entries;  // n * W (upper bound on nonzero digits)
  uint64_t max_chunks;   // ceil(max_entries / L)
  uint64_t tiles;        // sum over members of ceil(n_j / 4096) (sort tiles per window)
};

// Optional HIP-event marks recorded on the launch stream between pipeline
// stages, so callers can read per-stage device time of the last call.
struct StageTimer {
  static constexpr int N = 8;
  hipEvent_t ev[N] = {};
  bool used[N] = {};
  void mark(int i, hipStream_t st) {
    if (i < 0 || i >= N) return;
    if (!ev[i] && hipEventCreate(&ev[i]) != hipSuccess) return;
    if (hipEventRecord(ev[i], st) == hipSuccess) used[i] = true;
  }
  // ms between marks a and b (waits for b); -1 if either is missing
  float between(int a, int b) {
    if (!used[a] || !used[b]) return -1.f;
    if (hipEventSynchronize(ev[b]) != hipSuccess) return -1.f;
    float ms = -1.f;
    if (hipEventElapsedTime(&ms, ev[a], ev[b]) != hipSuccess) return -1.f;
    return ms;
  }
  void reset() {
    for (int i = 0; i < N; i++) used[i] = false;
  }
  void destroy() {
    for (int i = 0; i < N; i++)
      if (ev[i]) (void)hipEventDestroy(ev[i]);
  }
};

}  // namespace csk

struct csk_ctx {
  int device = 0;
  csk::Workspace ws;
  csk::Workspace ws2;  // second workspace for nested users (prover stages)
  csk::Workspace batch_ws;  // shared by batches created with CSK_BATCH_CTX_WORKSPACE
  std::mutex mu;       // serialises host-side use of the workspace
  uint32_t *g1_table = nullptr;  // fixed-base tables (k*2^(8w)*G), lazily built
  uint32_t *g2_table = nullptr;
  unsigned long long *check_res = nullptr;  // csk_ec_check's two counters
  bool timing = false;  // record stage events (csk_ctx_set_timing)
  csk::StageTimer timer;
  csk::MsmPlan last_plan = {};  // plan of the last multiexp (for reporting)
  csk_msm_cfg tune[2][64] = {};  // per (curve, floor(log2 n)) settings; zero = built-in choice
};
