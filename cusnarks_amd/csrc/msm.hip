// Pippenger multi-exponentiation for BN254 G1 / G2 on gfx950.
//
// Replaces the reference's Straus-8 GPU multiexp (cusnarks
// src/cuda/ecbn128_device.cu:508-565, 865-969, 1692-1866) and its CPU twin
// (ec_jacreduce_server_h, src/cuda/ec.cpp:1536-1704).  Both compute the same
// group element sum_i k_i P_i; the output here is the canonical affine point,
// so it is bit-identical to the reference's ec_jac2aff_h of its result.
//
// Pipeline (all on one stream, no host round trip):
//   1. k_digits      one thread per point: k mod r, signed c-bit digits for
//                    W windows; (window, |digit|) -> bucket key; histogram.
//   2. scan          exclusive prefix sum of the histogram -> bucket offsets.
//   3. k_scatter     counting-sort placement of (point index | sign) by key.
//   4. k_accumulate  each thread folds a fixed-length chunk of the sorted
//                    array into XYZZ bucket sums with mixed additions; runs
//                    that start in an earlier chunk are handed over as carries.
//   5. k_fixup       the owner of each bucket split over chunks adds the carries.
//   6. k_bucket_reduce  per (window, segment): sum_b (b+1) B_b by running sums.
//   7. k_window_reduce  per window: sum of the segment results.
//   8. k_final       Horner over the windows (c doublings each) -> affine.
#include <vector>
#include "internal.h"
#include "field.h"
#include "msm.h"

namespace csk {

static constexpr int SCAN_TILE = 256 * 8;  // matches msm_impl.h

static uint32_t floor_log2(uint64_t n) {
  uint32_t lg = 0;
  while ((1ull << (lg + 1)) <= n) lg++;
  return lg;
}

// ------------------------------------------------------------------ host side
uint32_t msm_choose_c(uint64_t n, bool g2, const csk_ctx *ctx, const csk_msm_cfg *cfg) {
  if (cfg && cfg->window_bits) return cfg->window_bits;
  const uint32_t lg = floor_log2(n);
  if (ctx && lg < 64 && ctx->tune[g2][lg].window_bits) return ctx->tune[g2][lg].window_bits;
  // windows whose top window keeps most of its bits (254 = (W-1) c + t with
  // t close to c): a narrow top window puts every point in a few buckets
  uint32_t c;
  if (lg <= 9) c = 8;
  else if (lg <= 13) c = 10;
  else if (lg <= 17) c = 13;
  else if (lg <= 19) c = 15;
  else c = 16;
  if (g2 && c > 15) c = 15;
  return c;
}

// Bits of the bucket index the sort's first pass bins by: the top 8 (at most
// c - 1), fewer for small members, so a second-pass workgroup (one per member,
// window and first-pass bin) gets about 2048 entries rather than a handful.
// The second pass keeps 3 * 2^(c-1-hb) counters in LDS, so hb >= c - 12.
static uint32_t plan_hi_bits(uint32_t c, uint64_t n_avg) {
  uint32_t hb = c - 1 < 8 ? c - 1 : 8;
  int ideal = (int)floor_log2(n_avg ? n_avg : 1) - 11;
  const int lo = (int)c - 12;
  if (ideal < lo) ideal = lo;
  if (ideal < 0) ideal = 0;
  if (ideal < (int)hb) hb = (uint32_t)ideal;
  return hb;
}

// Low bucket bits the second pass sorts the top window by.  That window
// holds only the scalar's top t = 254 mod c bits (scalars are reduced below
// r < 2^254, and its digit, at most 2^t <= 2^(c-1), is never negated), so its
// buckets are below 2^t: its first pass bins by the top hb of those t bits
// (all HB bins filled, as for the other windows) or, when t <= hb, by all of
// them, and each of its first-pass bins is then a single bucket.  Binning it
// like the others instead leaves most of its entries in a few bins, which the
// second pass takes one workgroup each for (c = 18 at 2^24 points: 22 ms).
static uint32_t plan_top_bits(uint32_t c, uint32_t hb) {
  const uint32_t t = 254 % c;
  return t > hb ? t - hb : 0;
}

MsmPlan msm_group_plan(uint32_t c, const uint64_t *n, uint32_t J, bool g2, uint32_t chunk, uint32_t reduce_seg) {
  (void)g2;
  MsmPlan pl = {};
  pl.c = c;
  pl.W = 254 / c + 1;
  pl.B = 1u << (c - 1);
  pl.K = pl.W * pl.B;
  pl.J = J;
  for (uint32_t j = 0; j < J; j++) {
    pl.n += n[j];
    pl.tiles += (n[j] + kSortTile - 1) / kSortTile;
  }
  pl.max_entries = pl.n * pl.W;
  pl.hb = plan_hi_bits(c, J ? pl.n / J : 0);
  pl.st = plan_top_bits(c, pl.hb);
  // chunk length: aim for ~2^18 accumulation threads (a few waves per SIMD)
  // so small inputs are not latency bound; at most 256 entries per thread
  uint64_t l = pl.max_entries >> 18;
  if (l < 8) l = 8;
  if (l > 256) l = 256;
  pl.L = chunk ? chunk : (uint32_t)l;
  // buckets per reduction thread: aim for ~2^15 reduction threads
  const uint64_t Ktot = (uint64_t)J * pl.K;
  uint32_t lr = 16;
  if (reduce_seg) lr = reduce_seg;
  else
    while (lr > 2 && Ktot / lr < (1u << 15)) lr >>= 1;
  if (lr > pl.B) lr = pl.B;
  pl.Lr = lr;
  pl.S = pl.B / pl.Lr;
  pl.max_chunks = (pl.max_entries + pl.L - 1) / pl.L;
  return pl;
}

bool msm_plan_ok(const MsmPlan &pl) {
  if (pl.c < 2 || pl.c > kMaxWindowBits || pl.J == 0) return false;
  if (pl.hb > 8 || pl.hb + 1 > pl.c || pl.c - 1 - pl.hb > 11) return false;
  if (pl.st > pl.c - 1 - pl.hb) return false;
  if (pl.Lr == 0 || (pl.Lr & (pl.Lr - 1)) != 0 || pl.B % pl.Lr) return false;
  if (pl.max_entries >= (1ull << 31)) return false;
  const uint32_t hi_bits = pl.hb;
  if ((uint64_t)pl.W * pl.tiles * (1u << hi_bits) >= (1ull << 32) - 1) return false;
  if ((uint64_t)pl.J * pl.K >= (1ull << 32) - 1) return false;
  if (pl.L == 0 || pl.L > (1u << 20)) return false;
  return true;
}

size_t msm_workspace_bytes(const MsmPlan &pl, bool g2) {
  // must mirror the carve order of msm_run (msm_impl.h)
  const uint32_t hi_bits = pl.hb;
  const uint64_t ncnt = (uint64_t)pl.W * (1u << hi_bits) * pl.tiles;
  const uint64_t nb = (ncnt + SCAN_TILE - 1) / SCAN_TILE;
  const uint64_t Ktot = (uint64_t)pl.J * pl.K, Wtot = (uint64_t)pl.J * pl.W;
  size_t b = 0;
  b += Carver::round(pl.max_entries * 4);       // digits
  b += Carver::round((pl.max_entries + 4) * 4); // sorted (+ 4: quad loads)
  b += Carver::round(msm_sortbuf_bytes(pl, g2)); // pass-1 entries, then 29-bit points and slots
  b += Carver::round((ncnt + 1) * 4);           // tile counts -> offsets
  b += Carver::round(nb * 4);                   // scan block sums
  b += Carver::round((Ktot + 1) * 4);           // bucket offsets
  b += Carver::round(pl.max_chunks * 4);        // carry keys
  b += Carver::round(Wtot * pl.S * (g2 ? 288 : 144)); // partials (29-bit slots)
  b += Carver::round(Wtot * (g2 ? 288 : 144));  // window sums (29-bit slots)
  return b + 4096;
}

void msm_fill_segs(const MsmPlan &pl, const uint64_t *n, const uint32_t *const *scalars,
                   const uint32_t *const *points, uint32_t *const *out, MsmSeg *segs) {
  const uint32_t hi_bits = pl.hb;
  const uint32_t HB = 1u << hi_bits;
  uint64_t pbase = 0, tiles = 0;
  for (uint32_t j = 0; j <= pl.J; j++) {
    MsmSeg &g = segs[j];
    g = MsmSeg{};
    g.pbase = pbase;
    g.tbase = (uint32_t)(pl.W * tiles);
    g.cbase = (uint32_t)(pl.W * HB * tiles);
    if (j == pl.J) break;
    g.scalars = scalars[j];
    g.points = points[j];
    g.out = out[j];
    g.n = n[j];
    g.NT = (uint32_t)((n[j] + kSortTile - 1) / kSortTile);
    g.qbase = (uint32_t)pbase;
    pbase += n[j];
    tiles += g.NT;
  }
}

// writes a single call's descriptor table (member + sentinel) on the stream
__global__ void k_put_segs(MsmSeg a, MsmSeg b, MsmSeg *dst) {
  if (threadIdx.x == 0) {
    dst[0] = a;
    dst[1] = b;
  }
}

int msm_launch(csk_ctx *ctx, bool g2, const uint32_t *d_scalars, const uint32_t *d_points, uint64_t n,
               uint32_t *d_out, const csk_msm_cfg *cfg, hipStream_t st) {
  if (!d_out) {
    set_error("msm: null output");
    return CSK_ERR_ARG;
  }
  CSK_HIP(hipSetDevice(ctx->device));
  if (n == 0) {
    // empty sum: the point at infinity, (0, 1*R)
    uint32_t inf[32] = {0};
    for (int i = 0; i < 8; i++) inf[(g2 ? 16 : 8) + i] = FieldConst<0>::one[i];
    CSK_HIP(hipMemcpyAsync(d_out, inf, (g2 ? 32 : 16) * 4, hipMemcpyHostToDevice, st));
    CSK_HIP(hipStreamSynchronize(st));
    return CSK_OK;
  }
  if (!d_scalars || !d_points) {
    set_error("msm: null input");
    return CSK_ERR_ARG;
  }
  const uint32_t c = msm_choose_c(n, g2, ctx, cfg);
  const uint32_t lg = floor_log2(n);
  const csk_msm_cfg *tn = lg < 64 ? &ctx->tune[g2][lg] : nullptr;
  const uint32_t chunk = cfg && cfg->chunk ? cfg->chunk : (tn && !(cfg && cfg->window_bits) ? tn->chunk : 0);
  const uint32_t rseg = cfg && cfg->reduce_seg ? cfg->reduce_seg : (tn && !(cfg && cfg->window_bits) ? tn->reduce_seg : 0);
  MsmPlan pl = msm_group_plan(c, &n, 1, g2, chunk, rseg);
  if (!msm_plan_ok(pl)) {
    set_error("msm: unsupported configuration (window/chunk/segment or size too large)");
    return CSK_ERR_ARG;
  }
  const size_t seg_bytes = Carver::round(2 * sizeof(MsmSeg));
  int rc = ctx->ws.ensure(seg_bytes + msm_workspace_bytes(pl, g2));
  if (rc) return rc;
  StageTimer *tm = nullptr;
  if (ctx->timing) {
    ctx->timer.reset();
    tm = &ctx->timer;
  }
  ctx->last_plan = pl;
  MsmSeg segs[2];
  msm_fill_segs(pl, &n, &d_scalars, &d_points, &d_out, segs);
  MsmSeg *d_segs = (MsmSeg *)ctx->ws.ptr;
  k_put_segs<<<1, 64, 0, st>>>(segs[0], segs[1], d_segs);
  CSK_CHECK_LAUNCH();
  void *ws = (char *)ctx->ws.ptr + seg_bytes;
  return g2 ? msm_run_g2(ws, pl, d_segs, st, tm) : msm_run_g1(ws, pl, d_segs, st, tm);
}

// ------------------------------------------------- multiexp without allocation
// csk_msm_g1_noalloc runs the same pipeline as csk_msm_g1 in a fixed device
// array instead of the context's growable workspace.  Every GPU that loads
// the library holds one copy of it (kNoAllocBytes) until the process ends;
// host code reaches it only through the address hipGetSymbolAddress gives
// for the current device, looked up once per device at context creation (or
// on the entry point's first call on that device).
static constexpr size_t kNoAllocBytes = 48ull << 20;
static constexpr uint64_t kNoAllocMaxPoints = 1ull << 15;
static constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// with `device` current: look up this device's copy of the array
int msm_noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("msm_noalloc: device index beyond the supported 64");
    return CSK_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return CSK_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error(std::string("msm_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return CSK_ERR_HIP;
  }
  g_scratch_ptr[device] = p;
  return CSK_OK;
}

// out = (0, 1*R), the affine point at infinity of the empty sum
__global__ void k_put_inf_g1(uint32_t *out) {
  uint32_t i = threadIdx.x;
  if (i < 16) out[i] = i < 8 ? 0u : FieldConst<0>::one[i - 8];
}

}  // namespace csk

using namespace csk;

extern "C" int csk_msm_g1_noalloc(csk_ctx *ctx, const uint32_t *d_scalars, const uint32_t *d_points, uint64_t n,
                                  uint32_t *d_out, const csk_msm_cfg *cfg, void *stream) {
  if (!ctx || !d_out) {
    set_error("msm_g1_noalloc: null context or output");
    return CSK_ERR_ARG;
  }
  if (n > kNoAllocMaxPoints) {
    set_error("msm_g1_noalloc: n = " + std::to_string(n) + " exceeds the no-allocation limit of " +
              std::to_string(kNoAllocMaxPoints) + " points; use csk_msm_g1");
    return CSK_ERR_ARG;
  }
  if (n && (!d_scalars || !d_points)) {
    set_error("msm_g1_noalloc: null input");
    return CSK_ERR_ARG;
  }
  CSK_HIP(hipSetDevice(ctx->device));
  int rc = msm_noalloc_setup(ctx->device);
  if (rc) return rc;
  char *base = (char *)g_scratch_ptr[ctx->device];
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) {
    k_put_inf_g1<<<1, 64, 0, st>>>(d_out);
    CSK_CHECK_LAUNCH();
    return CSK_OK;
  }
  // the tuned window width, narrowed until the workspace fits the array
  // (every width gives the same sum); an explicit width that does not fit is an error
  const bool fixed_c = cfg && cfg->window_bits;
  uint32_t c = msm_choose_c(n, false, ctx, cfg);
  const uint32_t lg = floor_log2(n);
  const csk_msm_cfg *tn = &ctx->tune[0][lg];
  const uint32_t chunk = cfg && cfg->chunk ? cfg->chunk : (!fixed_c ? tn->chunk : 0);
  const uint32_t rseg = cfg && cfg->reduce_seg ? cfg->reduce_seg : (!fixed_c ? tn->reduce_seg : 0);
  const size_t seg_bytes = Carver::round(2 * sizeof(MsmSeg));
  MsmPlan pl = msm_group_plan(c, &n, 1, false, chunk, rseg);
  while (!fixed_c && c > 2 && (!msm_plan_ok(pl) || seg_bytes + msm_workspace_bytes(pl, false) > kNoAllocBytes)) {
    c--;
    pl = msm_group_plan(c, &n, 1, false, chunk, rseg);
  }
  if (!msm_plan_ok(pl)) {
    set_error("msm_g1_noalloc: unsupported configuration (window/chunk/segment)");
    return CSK_ERR_ARG;
  }
  if (seg_bytes + msm_workspace_bytes(pl, false) > kNoAllocBytes) {
    set_error("msm_g1_noalloc: the workspace for window width " + std::to_string(pl.c) + " needs " +
              std::to_string(seg_bytes + msm_workspace_bytes(pl, false)) + " bytes, more than the " +
              std::to_string(kNoAllocBytes) + "-byte scratch array");
    return CSK_ERR_ARG;
  }
  ctx->last_plan = pl;
  MsmSeg segs[2];
  msm_fill_segs(pl, &n, &d_scalars, &d_points, &d_out, segs);
  MsmSeg *d_segs = (MsmSeg *)base;
  k_put_segs<<<1, 64, 0, st>>>(segs[0], segs[1], d_segs);
  CSK_CHECK_LAUNCH();
  return msm_run_g1(base + seg_bytes, pl, d_segs, st, nullptr);
}

// ------------------------------------------------------------ fixed-base tables
// csk_msm_table_*: the table multiexp of msm_impl.h (window-shifted copies of
// fixed points, one bucket reduction, no doublings between windows).
struct csk_msm_table {
  csk_ctx *ctx = nullptr;
  int ec2 = 0;
  uint64_t n = 0, n_pad = 0;
  uint32_t c = 0, Wt = 0;
  uint32_t *table = nullptr;  // Wt * n_pad points in the 2^261 form
  uint32_t *live = nullptr;   // one bit per point: not at infinity
  uint32_t Lrm = 0, Sm = 0;
  uint64_t table_bytes = 0;
};

namespace csk {

// group plan of the count * Wt one-window members of runs over count scalar
// vectors on a table of n_pad points
static MsmPlan table_plan(uint32_t c, uint64_t n_pad, uint32_t Wt, uint32_t count) {
  MsmPlan pl = {};
  pl.c = c;
  pl.W = 1;
  pl.B = 1u << (c - 1);
  pl.K = pl.B;
  pl.J = Wt * count;
  pl.n = n_pad * pl.J;
  pl.tiles = pl.J * (n_pad / kSortTile);
  pl.max_entries = pl.n;
  pl.hb = plan_hi_bits(c, n_pad);
  pl.st = plan_top_bits(c, pl.hb);
  uint64_t l = pl.max_entries >> 18;
  if (l < 8) l = 8;
  if (l > 256) l = 256;
  pl.L = (uint32_t)l;
  pl.Lr = pl.B < 16 ? pl.B : 16;
  pl.S = pl.B / pl.Lr;
  pl.max_chunks = (pl.max_entries + pl.L - 1) / pl.L;
  return pl;
}

static uint64_t table_point_bytes(int ec2, uint64_t n_pad, uint32_t Wt) {
  return (uint64_t)Wt * n_pad * (ec2 ? 128 : 64) + n_pad / 8;
}

// A table run reduces one set of 2^(c-1) buckets whatever Wt is, so it
// affords wider windows than the plain multiexp: 16 from 2^15 points, 17
// from 2^18 and 20 (13 windows instead of 15) from 3 * 2^22 G1 and 2^24 G2
// points (measured on G1 and G2 at 2^16 .. 2^25: G1 2^24 20.8 against 21.6 ms
// at c = 17, 2^25 38.9 against 41.7, G2 2^24 54.4 against 55.5; G1 2^23 ties,
// G2 2^23 33.4 against 30.5), the plain width below.  A width whose top
// window keeps few scalar bits (t = 254 mod c: 2 at c = 14 and 18) puts that
// window's every entry into 2^t buckets, whose long carry chains cost more
// than the window it saves (2^14 points: 1.09 ms at c = 14, 0.76 ms at 15):
// 14 becomes 15.
static uint32_t table_default_c(uint64_t n, int ec2, const csk_ctx *ctx) {
  const uint32_t lg = floor_log2(n);
  if (n >= (ec2 ? 1ull << 24 : 3ull << 22)) return 20;
  if (lg >= 18) return 17;
  if (lg >= 15) return 16;
  uint32_t c = msm_choose_c(n, ec2 != 0, ctx, nullptr);
  if (c > 16) c = 16;
  if (c == 14) c = 15;
  return c;
}

// count runs in one pipeline, run p over table t[p] (same curve and window
// width), in the workspace of ctx, on stream st
static int table_run(csk_ctx *ctx, uint32_t count, csk_msm_table *const *t, const uint32_t *const *d_scalars,
                     uint32_t *const *d_out, hipStream_t st) {
  const csk_msm_table *t0 = t[0];
  const uint32_t Wt = t0->Wt, c = t0->c;
  MsmPlan pl = table_plan(c, t0->n_pad, Wt, 1);  // B, K, hb, L are set below for the whole group
  pl.J = count * Wt;
  pl.n = 0;
  pl.tiles = 0;
  uint32_t max_nt = 0;
  for (uint32_t p = 0; p < count; p++) {
    if (t[p]->ec2 != t0->ec2 || t[p]->c != c || t[p]->ctx->device != ctx->device) {
      set_error("msm_table_run: the tables of one run must share curve, window width and device");
      return CSK_ERR_ARG;
    }
    const uint32_t nt = (uint32_t)(t[p]->n_pad / kSortTile);
    pl.n += (uint64_t)Wt * t[p]->n_pad;
    pl.tiles += (uint64_t)Wt * nt;
    if (nt > max_nt) max_nt = nt;
  }
  pl.max_entries = pl.n;
  pl.hb = plan_hi_bits(c, pl.n / pl.J);
  pl.st = plan_top_bits(c, pl.hb);
  uint64_t l = pl.max_entries >> 18;
  pl.L = (uint32_t)(l < 8 ? 8 : l > 256 ? 256 : l);
  pl.max_chunks = (pl.max_entries + pl.L - 1) / pl.L;
  if (!msm_plan_ok(pl)) {
    set_error("msm_table_run: " + std::to_string(count) + " runs of up to " + std::to_string(max_nt * kSortTile) +
              " points exceed the multiexp's index range");
    return CSK_ERR_ARG;
  }
  CSK_HIP(hipSetDevice(ctx->device));
  const uint64_t PW = t0->ec2 ? 32 : 16;
  std::vector<uint64_t> ns(pl.J);
  std::vector<const uint32_t *> sc(pl.J, nullptr), pts(pl.J);
  std::vector<uint32_t *> outs(pl.J, nullptr);
  for (uint32_t m = 0; m < pl.J; m++) {
    const csk_msm_table *tp = t[m / Wt];
    ns[m] = tp->n_pad;
    pts[m] = tp->table + (uint64_t)(m % Wt) * tp->n_pad * PW;
  }
  std::vector<MsmSeg> segs(pl.J + 1);
  msm_fill_segs(pl, ns.data(), sc.data(), pts.data(), outs.data(), segs.data());
  int rc = ctx->ws.ensure(msm_table_ws_bytes(pl, t0->ec2 != 0, t0->Sm, count));
  if (rc) return rc;
  TabArgs args = {};
  for (uint32_t p = 0; p < count; p++) {
    args.s[p] = d_scalars[p];
    args.o[p] = d_out[p];
    args.live[p] = t[p]->live;
    args.n[p] = t[p]->n;
  }
  return t0->ec2 ? msm_run_table_g2(ctx->ws.ptr, pl, segs.data(), args, count, max_nt, t0->Lrm, t0->Sm, st)
                 : msm_run_table_g1(ctx->ws.ptr, pl, segs.data(), args, count, max_nt, t0->Lrm, t0->Sm, st);
}

}  // namespace csk

extern "C" int csk_msm_table_bytes(int ec2, uint64_t n, uint32_t window_bits, uint64_t *bytes) {
  if (!bytes || n == 0 || window_bits < 2 || window_bits > kMaxWindowBits) {
    set_error("msm_table_bytes: n >= 1 and 2 <= window_bits <= 20 required");
    return CSK_ERR_ARG;
  }
  const uint64_t n_pad = (n + kSortTile - 1) / kSortTile * kSortTile;
  *bytes = table_point_bytes(ec2, n_pad, 254 / window_bits + 1);
  return CSK_OK;
}

extern "C" int csk_msm_table_create(csk_ctx *ctx, int ec2, const uint32_t *d_points, uint64_t n, uint32_t window_bits,
                                    csk_msm_table **out, void *stream) {
  if (!ctx || !out || !d_points || n == 0) {
    set_error("msm_table_create: null argument or n == 0");
    return CSK_ERR_ARG;
  }
  *out = nullptr;
  const uint32_t c = window_bits ? window_bits : table_default_c(n, ec2, ctx);
  if (c < 2 || c > kMaxWindowBits) {
    set_error("msm_table_create: window_bits must be 2..20");
    return CSK_ERR_ARG;
  }
  const uint32_t Wt = 254 / c + 1;
  const uint64_t n_pad = (n + kSortTile - 1) / kSortTile * kSortTile;
  MsmPlan pl = table_plan(c, n_pad, Wt, 1);
  if (!msm_plan_ok(pl)) {
    set_error("msm_table_create: " + std::to_string(n) + " points x " + std::to_string(Wt) +
              " windows exceed the multiexp's index range");
    return CSK_ERR_ARG;
  }
  CSK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  csk_msm_table *t = new csk_msm_table();
  t->ctx = ctx;
  t->ec2 = ec2 ? 1 : 0;
  t->n = n;
  t->n_pad = n_pad;
  t->c = c;
  t->Wt = Wt;
  uint32_t lr = 16;
  while (lr > 2 && pl.B / lr < (1u << 15)) lr >>= 1;
  if (lr > pl.B) lr = pl.B;
  t->Lrm = lr;
  t->Sm = pl.B / lr;
  const uint64_t PW = ec2 ? 32 : 16;
  t->table_bytes = table_point_bytes(t->ec2, n_pad, Wt);
  hipError_t e = hipMalloc(&t->table, (size_t)Wt * n_pad * PW * 4);
  if (e == hipSuccess) e = hipMalloc(&t->live, (size_t)(n_pad / 32) * 4);
  if (e != hipSuccess) {
    set_error(std::string("msm_table_create: allocation: ") + hipGetErrorString(e));
    csk_msm_table_destroy(t);
    return CSK_ERR_NOMEM;
  }
  int rc = t->ec2 ? msm_table_build_g2(d_points, n, n_pad, c, Wt, t->table, t->live, st)
                  : msm_table_build_g1(d_points, n, n_pad, c, Wt, t->table, t->live, st);
  if (rc == CSK_OK) {
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      set_error(std::string("msm_table_create: ") + hipGetErrorString(e));
      rc = CSK_ERR_HIP;
    }
  }
  if (rc) {
    csk_msm_table_destroy(t);
    return rc;
  }
  *out = t;
  return CSK_OK;
}

extern "C" int csk_msm_table_run(csk_msm_table *t, const uint32_t *d_scalars, uint32_t *d_out, void *stream) {
  if (!t || !d_scalars || !d_out) {
    set_error("msm_table_run: null argument");
    return CSK_ERR_ARG;
  }
  return table_run(t->ctx, 1, &t, &d_scalars, &d_out, (hipStream_t)stream);
}

extern "C" int csk_msm_table_run_batch(csk_ctx *ctx, csk_msm_table *t, uint32_t count,
                                       const uint32_t *const *d_scalars, uint32_t *const *d_out, void *stream) {
  if (!t || (count && (!d_scalars || !d_out)) || count > kTabBatch) {
    set_error("msm_table_run_batch: null argument or more than 64 scalar vectors");
    return CSK_ERR_ARG;
  }
  if (!ctx) ctx = t->ctx;
  if (ctx->device != t->ctx->device) {
    set_error("msm_table_run_batch: the context's device is not the table's");
    return CSK_ERR_ARG;
  }
  for (uint32_t p = 0; p < count; p++)
    if (!d_scalars[p] || !d_out[p]) {
      set_error("msm_table_run_batch: null buffer for scalar vector " + std::to_string(p));
      return CSK_ERR_ARG;
    }
  if (count == 0) return CSK_OK;
  std::vector<csk_msm_table *> ts(count, t);
  return table_run(ctx, count, ts.data(), d_scalars, d_out, (hipStream_t)stream);
}

extern "C" int csk_msm_table_run_multi(csk_ctx *ctx, uint32_t count, csk_msm_table *const *tables,
                                       const uint32_t *const *d_scalars, uint32_t *const *d_out, void *stream) {
  if (!ctx || (count && (!tables || !d_scalars || !d_out)) || count > kTabBatch) {
    set_error("msm_table_run_multi: null argument or more than 64 runs");
    return CSK_ERR_ARG;
  }
  for (uint32_t p = 0; p < count; p++)
    if (!tables[p] || !d_scalars[p] || !d_out[p]) {
      set_error("msm_table_run_multi: null table or buffer for run " + std::to_string(p));
      return CSK_ERR_ARG;
    }
  if (count == 0) return CSK_OK;
  return table_run(ctx, count, tables, d_scalars, d_out, (hipStream_t)stream);
}

extern "C" int csk_msm_table_info(const csk_msm_table *t, uint64_t *out, int n) {
  if (!t || !out || n < 1) return CSK_ERR_ARG;
  const MsmPlan pl = table_plan(t->c, t->n_pad, t->Wt, 1);
  const uint64_t v[5] = {t->n, t->c, t->Wt, t->table_bytes, msm_table_ws_bytes(pl, t->ec2 != 0, t->Sm, 1)};
  for (int i = 0; i < n && i < 5; i++) out[i] = v[i];
  return CSK_OK;
}

extern "C" void csk_msm_table_destroy(csk_msm_table *t) {
  if (!t) return;
  if (t->ctx) (void)hipSetDevice(t->ctx->device);
  if (t->table || t->live) (void)hipDeviceSynchronize();
  if (t->table) (void)hipFree(t->table);
  if (t->live) (void)hipFree(t->live);
  delete t;
}
