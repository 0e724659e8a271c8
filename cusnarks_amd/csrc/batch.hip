// Batched multi-exponentiation: many uneven multiexps in one pipeline.
//
// The reference splits every prover multiexp into chunks of at most 2^20
// points and deals them round-robin over (GPU, stream) pairs, chaining each
// chunk's partial result through a (scalar 1, previous point) slot and adding
// the per-stream results on the CPU (findECPointsDispatch / init_EC_P,
// cusnarks src/python/groth_prover.py:1548-1688; buildDispatchTable,
// src/python/cuda_wrapper.py:1643-1677).  Here a batch is planned once:
//
//   * members are grouped by window width c (per size class: tuned table or
//     the built-in choice); a group whose entry count would pass the 32-bit
//     entry index space is split;
//   * a group runs the ordinary Pippenger pipeline over the concatenated key
//     space (member j owns keys [j*K, (j+1)*K)), so each stage is one launch
//     over all members and small members no longer pay a whole pipeline of
//     launches and a serial window combination each;
//   * groups are dealt over up to four streams forked from the caller's
//     stream and joined back into it;
//   * the enqueue sequence is free of allocation, host synchronisation and
//     memset nodes, so it is captured once into a HIP graph; member pointers
//     live in device descriptor tables, so a replay may use new buffers.
#include <algorithm>
#include <cstring>
#include <vector>
#include "internal.h"
#include "field.h"
#include "msm.h"

using namespace csk;

struct csk_msm_batch {
  struct Group {
    MsmPlan pl;
    std::vector<uint32_t> members;  // batch indices, ascending
    std::vector<MsmSeg> host_segs;  // last uploaded table (J + 1)
    MsmSeg *d_segs = nullptr;
    void *ws = nullptr;
    int stream = 0;  // index into streams
  };
  csk_ctx *ctx = nullptr;
  bool g2 = false;
  uint32_t count = 0;
  std::vector<uint64_t> n;
  std::vector<Group> groups;
  void *mem = nullptr;    // descriptor tables (+ workspaces unless shared)
  size_t bytes = 0;       // device bytes the batch needs (tables + workspaces)
  size_t ws_bytes = 0;    // workspace part
  bool shared = false;    // workspaces live in ctx->batch_ws
  uint64_t ws_gen = 0;    // generation of ctx->batch_ws the pointers were carved from
  std::vector<size_t> ws_off;  // per group, offset of its workspace
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr, done = nullptr;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ran = false;
};

namespace {

constexpr uint64_t kGroupEntryCap = 1ull << 30;  // entries per group (16 B of workspace each)
constexpr int kMaxStreams = 4;

void destroy_batch(csk_msm_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  if (b->ran) (void)hipEventSynchronize(b->done);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (hipEvent_t e : b->joins)
    if (e) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->done) (void)hipEventDestroy(b->done);
  for (hipStream_t s : b->streams)
    if (s) (void)hipStreamDestroy(s);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  if (b->mem) (void)hipFree(b->mem);
  delete b;
}

// fork from `origin`, enqueue every group on its stream, join back
int enqueue(csk_msm_batch *b, hipStream_t origin) {
  CSK_HIP(hipEventRecord(b->fork, origin));
  for (hipStream_t s : b->streams) CSK_HIP(hipStreamWaitEvent(s, b->fork, 0));
  for (auto &g : b->groups) {
    hipStream_t s = b->streams[g.stream];
    int rc = b->g2 ? msm_run_g2(g.ws, g.pl, g.d_segs, s, nullptr) : msm_run_g1(g.ws, g.pl, g.d_segs, s, nullptr);
    if (rc) return rc;
  }
  for (size_t i = 0; i < b->streams.size(); i++) {
    CSK_HIP(hipEventRecord(b->joins[i], b->streams[i]));
    CSK_HIP(hipStreamWaitEvent(origin, b->joins[i], 0));
  }
  return CSK_OK;
}

int capture(csk_msm_batch *b) {
  CSK_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
  int rc = enqueue(b, b->cap);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(b->cap, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) {
    set_error(std::string("msm batch capture: ") + hipGetErrorString(e));
    return CSK_ERR_HIP;
  }
  b->graph = graph;
  CSK_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return CSK_OK;
}

}  // namespace

extern "C" {

int csk_msm_set_tuning(csk_ctx *ctx, int ec2, const csk_msm_cfg *table, uint32_t count) {
  if (!ctx || (count && !table) || count > 64) {
    set_error("msm_set_tuning: bad arguments");
    return CSK_ERR_ARG;
  }
  const bool g2 = ec2 != 0;
  for (uint32_t lg = 0; lg < 64; lg++) {
    csk_msm_cfg e = lg < count ? table[lg] : csk_msm_cfg{};
    if (e.window_bits && (e.window_bits < 2 || e.window_bits > kMaxWindowBits)) {
      set_error("msm_set_tuning: window_bits out of range");
      return CSK_ERR_ARG;
    }
    if (e.reduce_seg && (e.reduce_seg & (e.reduce_seg - 1))) {
      set_error("msm_set_tuning: reduce_seg must be a power of two");
      return CSK_ERR_ARG;
    }
  }
  for (uint32_t lg = 0; lg < 64; lg++) ctx->tune[g2][lg] = lg < count ? table[lg] : csk_msm_cfg{};
  return CSK_OK;
}

}  // extern "C"

namespace {

// groups, plans and streams of a batch (no device memory yet)
int plan_batch(csk_msm_batch *b, const uint64_t *n, const uint32_t *window_bits) {
  const uint32_t count = b->count;
  const bool g2 = b->g2;
  // window width per member, then groups of equal width in member order
  // Many members of one size (a cohort of proofs of one circuit) pay their
  // buckets' reduction, W * 2^(c-1) buckets each, beside n * W accumulations,
  // while the batch amortises the fixed costs that the single-call tuning
  // table weighs.  For a size class (floor(log2 n)) with 8 or more members
  // the width is therefore capped near the minimum of W (n + 2.8 * 2^(c-1)),
  // c ~ log2 n - 3, on a coarse set of widths (fewer groups): 8 up to 2^12
  // points, 11 up to 2^15, 13 up to 2^17, 15 up to 2^19.
  std::vector<uint32_t> cw(count), lgs(count), per_class(64, 0);
  for (uint32_t i = 0; i < count; i++) {
    lgs[i] = 63 - __builtin_clzll(n[i] ? n[i] : 1);
    if (!(window_bits && window_bits[i])) per_class[lgs[i]]++;
  }
  for (uint32_t i = 0; i < count; i++) {
    csk_msm_cfg cf = {};
    cf.window_bits = window_bits ? window_bits[i] : 0;
    cw[i] = msm_choose_c(n[i] ? n[i] : 1, g2, b->ctx, &cf);
    if (!cf.window_bits && per_class[lgs[i]] >= 8) {
      const uint32_t lg = lgs[i];
      const uint32_t cap = lg <= 12 ? 8 : lg <= 15 ? 11 : lg <= 17 ? 13 : lg <= 19 ? 15 : kMaxWindowBits;
      if (cw[i] > cap) cw[i] = cap;
    }
    if (cw[i] < 2 || cw[i] > kMaxWindowBits || n[i] * (254 / cw[i] + 1) >= (1ull << 31)) {
      set_error("msm_batch: member " + std::to_string(i) + " too large or bad window width");
      return CSK_ERR_ARG;
    }
  }
  std::vector<uint32_t> widths(cw);
  std::sort(widths.begin(), widths.end());
  widths.erase(std::unique(widths.begin(), widths.end()), widths.end());
  for (uint32_t c : widths) {
    const uint64_t W = 254 / c + 1;
    csk_msm_batch::Group cur;
    uint64_t entries = 0;
    for (uint32_t i = 0; i < count; i++) {
      if (cw[i] != c) continue;
      if (!cur.members.empty() && entries + n[i] * W > kGroupEntryCap) {
        b->groups.push_back(cur);
        cur = csk_msm_batch::Group();
        entries = 0;
      }
      cur.members.push_back(i);
      entries += n[i] * W;
    }
    if (!cur.members.empty()) b->groups.push_back(cur);
  }
  for (auto &g : b->groups) {
    std::vector<uint64_t> gn;
    for (uint32_t i : g.members) gn.push_back(n[i]);
    g.pl = msm_group_plan(cw[g.members[0]], gn.data(), (uint32_t)gn.size(), g2, 0, 0);
    if (!msm_plan_ok(g.pl)) {
      set_error("msm_batch: unsupported group plan");
      return CSK_ERR_ARG;
    }
  }
  // largest groups first, dealt round-robin over the streams
  std::stable_sort(b->groups.begin(), b->groups.end(),
                   [](const csk_msm_batch::Group &x, const csk_msm_batch::Group &y) {
                     return x.pl.max_entries > y.pl.max_entries;
                   });
  const int ns = std::min<int>(kMaxStreams, (int)b->groups.size());
  for (size_t gi = 0; gi < b->groups.size(); gi++) b->groups[gi].stream = (int)(gi % ns);
  b->streams.assign(ns, nullptr);
  b->joins.assign(ns, nullptr);
  // layout: descriptor tables, then (unless shared) the workspaces
  size_t tables = 0;
  for (auto &g : b->groups) tables += Carver::round((g.pl.J + 1) * sizeof(MsmSeg));
  b->ws_off.assign(b->groups.size(), 0);
  size_t ws = 0;
  for (size_t gi = 0; gi < b->groups.size(); gi++) {
    b->ws_off[gi] = ws;
    ws += Carver::round(msm_workspace_bytes(b->groups[gi].pl, g2));
  }
  b->ws_bytes = ws;
  b->bytes = tables + ws;
  return CSK_OK;
}

// point the groups at their workspaces (own memory or the context's)
int bind_workspaces(csk_msm_batch *b) {
  char *base;
  if (b->shared) {
    int rc = b->ctx->batch_ws.ensure(b->ws_bytes);
    if (rc) return rc;
    base = (char *)b->ctx->batch_ws.ptr;
    if (b->ws_gen != b->ctx->batch_ws.gen) {
      // the captured graph holds the old addresses
      if (b->exec) (void)hipGraphExecDestroy(b->exec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
      b->ws_gen = b->ctx->batch_ws.gen;
    }
  } else {
    base = (char *)b->mem + (b->bytes - b->ws_bytes);
  }
  for (size_t gi = 0; gi < b->groups.size(); gi++) b->groups[gi].ws = base + b->ws_off[gi];
  return CSK_OK;
}

}  // namespace

extern "C" {

int csk_msm_batch_bytes(csk_ctx *ctx, int ec2, uint32_t count, const uint64_t *n, const uint32_t *window_bits,
                        uint64_t *bytes) {
  if (!ctx || !count || !n || !bytes) {
    set_error("msm_batch_bytes: bad arguments");
    return CSK_ERR_ARG;
  }
  csk_msm_batch b;
  b.ctx = ctx;
  b.g2 = ec2 != 0;
  b.count = count;
  int rc = plan_batch(&b, n, window_bits);
  if (rc) return rc;
  *bytes = b.ws_bytes;
  return CSK_OK;
}

int csk_msm_batch_create(csk_ctx *ctx, int ec2, uint32_t count, const uint64_t *n,
                         const uint32_t *window_bits, uint32_t flags, csk_msm_batch **out) {
  if (!ctx || !out || !count || !n || (flags & ~CSK_BATCH_CTX_WORKSPACE)) {
    set_error("msm_batch_create: bad arguments");
    return CSK_ERR_ARG;
  }
  *out = nullptr;
  CSK_HIP(hipSetDevice(ctx->device));
  csk_msm_batch *b = new csk_msm_batch();
  b->ctx = ctx;
  b->g2 = ec2 != 0;
  b->count = count;
  b->n.assign(n, n + count);
  b->shared = (flags & CSK_BATCH_CTX_WORKSPACE) != 0;
  int rc = plan_batch(b, n, window_bits);
  if (rc) {
    destroy_batch(b);
    return rc;
  }
  const size_t own = b->shared ? b->bytes - b->ws_bytes : b->bytes;
  hipError_t e = hipMalloc(&b->mem, own);
  if (e != hipSuccess) {
    set_error(std::string("msm_batch_create: device memory (") + std::to_string(own >> 20) + " MiB): " +
              hipGetErrorString(e));
    b->mem = nullptr;
    destroy_batch(b);
    return CSK_ERR_NOMEM;
  }
  size_t off = 0;
  for (auto &g : b->groups) {
    g.d_segs = (MsmSeg *)((char *)b->mem + off);
    off += Carver::round((g.pl.J + 1) * sizeof(MsmSeg));
  }
  rc = bind_workspaces(b);
  if (rc) {
    destroy_batch(b);
    return rc;
  }
  hipError_t ok = hipSuccess;
  for (size_t i = 0; i < b->streams.size() && ok == hipSuccess; i++) {
    ok = hipStreamCreateWithFlags(&b->streams[i], hipStreamNonBlocking);
    if (ok == hipSuccess) ok = hipEventCreateWithFlags(&b->joins[i], hipEventDisableTiming);
  }
  if (ok == hipSuccess) ok = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (ok == hipSuccess) ok = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
  if (ok == hipSuccess) ok = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
  if (ok != hipSuccess) {
    set_error(std::string("msm_batch_create: ") + hipGetErrorString(ok));
    destroy_batch(b);
    return CSK_ERR_HIP;
  }
  *out = b;
  return CSK_OK;
}

int csk_msm_batch_run(csk_msm_batch *b, const uint32_t *const *d_scalars, const uint32_t *const *d_points,
                      uint32_t *const *d_out, int use_graph, void *stream) {
  if (!b || !d_scalars || !d_points || !d_out) {
    set_error("msm_batch_run: null argument");
    return CSK_ERR_ARG;
  }
  for (uint32_t i = 0; i < b->count; i++)
    if (!d_out[i] || (b->n[i] && (!d_scalars[i] || !d_points[i]))) {
      set_error("msm_batch_run: null pointer for member " + std::to_string(i));
      return CSK_ERR_ARG;
    }
  CSK_HIP(hipSetDevice(b->ctx->device));
  hipStream_t st = (hipStream_t)stream;
  // refresh the descriptor tables whose pointers changed (after the previous
  // run of this batch, which reads them, has finished)
  bool waited = false;
  for (auto &g : b->groups) {
    const uint32_t J = g.pl.J;
    std::vector<uint64_t> gn(J);
    std::vector<const uint32_t *> s(J), p(J);
    std::vector<uint32_t *> o(J);
    for (uint32_t k = 0; k < J; k++) {
      uint32_t i = g.members[k];
      gn[k] = b->n[i];
      s[k] = d_scalars[i];
      p[k] = d_points[i];
      o[k] = d_out[i];
    }
    std::vector<MsmSeg> segs(J + 1);
    msm_fill_segs(g.pl, gn.data(), s.data(), p.data(), o.data(), segs.data());
    if (g.host_segs.size() == segs.size() &&
        std::memcmp(g.host_segs.data(), segs.data(), segs.size() * sizeof(MsmSeg)) == 0)
      continue;
    if (b->ran && !waited) {
      CSK_HIP(hipEventSynchronize(b->done));
      waited = true;
    }
    CSK_HIP(hipMemcpy(g.d_segs, segs.data(), segs.size() * sizeof(MsmSeg), hipMemcpyHostToDevice));
    g.host_segs.swap(segs);
  }
  if (b->shared) {
    int rc = bind_workspaces(b);
    if (rc) return rc;
  }
  if (use_graph) {
    if (!b->exec) {
      int rc = capture(b);
      if (rc) return rc;
    }
    CSK_HIP(hipGraphLaunch(b->exec, st));
  } else {
    int rc = enqueue(b, st);
    if (rc) return rc;
  }
  CSK_HIP(hipEventRecord(b->done, st));
  b->ran = true;
  return CSK_OK;
}

int csk_msm_batch_info(const csk_msm_batch *b, uint64_t *out, int n) {
  if (!b || !out) return CSK_ERR_ARG;
  std::vector<uint64_t> v = {b->groups.size(), b->bytes};
  for (auto &g : b->groups) {
    v.push_back(g.pl.c);
    v.push_back(g.pl.J);
    v.push_back(g.pl.L);
    v.push_back(g.pl.max_entries);
  }
  for (int i = 0; i < n && i < (int)v.size(); i++) out[i] = v[i];
  return CSK_OK;
}

void csk_msm_batch_destroy(csk_msm_batch *b) { destroy_batch(b); }

}  // extern "C"
