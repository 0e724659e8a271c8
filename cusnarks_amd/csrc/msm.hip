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
#include "internal.h"
#include "field.h"
#include "msm.h"

namespace csk {

static constexpr int SCAN_TILE = 256 * 8;  // matches msm_impl.h

// ------------------------------------------------------------------ host side
MsmPlan msm_plan(uint64_t n, bool g2, const csk_msm_cfg *cfg) {
  MsmPlan pl;
  uint32_t lg = 0;
  while ((1ull << (lg + 1)) <= n) lg++;
  uint32_t c = cfg && cfg->window_bits ? cfg->window_bits : 0;
  if (!c) {
    // windows whose top window keeps most of its bits (254 = (W-1) c + t with
    // t close to c): a narrow top window puts every point in a few buckets
    if (lg <= 9) c = 8;
    else if (lg <= 13) c = 10;
    else if (lg <= 17) c = 13;
    else if (lg <= 19) c = 15;
    else c = 16;
    if (g2 && c > 15) c = 15;
  }
  pl.c = c;
  pl.W = 254 / c + 1;
  pl.B = 1u << (c - 1);
  pl.K = pl.W * pl.B;
  pl.n = n;
  pl.max_entries = n * pl.W;
  // chunk length: aim for ~2^18 accumulation threads (a few waves per SIMD)
  // so small inputs are not latency bound; at most 256 entries per thread
  uint64_t l = pl.max_entries >> 18;
  if (l < 8) l = 8;
  if (l > 256) l = 256;
  pl.L = cfg && cfg->chunk ? cfg->chunk : (uint32_t)l;
  // buckets per reduction thread: aim for ~2^15 reduction threads
  uint32_t lr = 16;
  if (!(cfg && cfg->reduce_seg)) {
    while (lr > 2 && (uint64_t)pl.K / lr < (1u << 15)) lr >>= 1;
  } else lr = cfg->reduce_seg;
  if (lr > pl.B) lr = pl.B;
  pl.Lr = lr;
  pl.S = pl.B / pl.Lr;
  pl.max_chunks = (pl.max_entries + pl.L - 1) / pl.L;
  return pl;
}

static bool plan_ok(const MsmPlan &pl) {
  if (pl.c < 2 || pl.c > 22) return false;
  if ((pl.Lr & (pl.Lr - 1)) != 0 || pl.B % pl.Lr) return false;
  if (pl.max_entries >= (1ull << 31)) return false;
  if ((uint64_t)pl.W * 256 * ((pl.n + 4095) / 4096) >= (1ull << 32) - 1) return false;
  if (pl.L == 0 || pl.L > (1u << 20)) return false;
  return true;
}

size_t msm_workspace_bytes(const MsmPlan &pl, bool g2) {
  // must mirror the carve order of msm_run (msm_impl.h)
  const size_t XB = (g2 ? 64 : 32) * 4;
  const uint64_t NT = (pl.n + 4095) / 4096;
  const uint32_t hi_bits = pl.c - 1 < 8 ? pl.c - 1 : 8;
  const uint64_t ncnt = (uint64_t)pl.W * (1u << hi_bits) * NT;
  const uint64_t nb = (ncnt + SCAN_TILE - 1) / SCAN_TILE;
  size_t b = 0;
  b += Carver::round(pl.max_entries * 4);       // digits
  b += Carver::round(pl.max_entries * 4);       // sorted
  b += Carver::round(pl.max_entries * 8);       // pass-1 entries
  b += Carver::round((ncnt + 1) * 4);           // tile counts -> offsets
  b += Carver::round(nb * 4);                   // scan block sums
  b += Carver::round(((size_t)pl.K + 1) * 4);   // bucket offsets
  b += Carver::round((size_t)pl.K * XB);        // buckets
  b += Carver::round(pl.max_chunks * XB);       // carries
  b += Carver::round(pl.max_chunks * 4);        // carry keys
  b += Carver::round((size_t)pl.W * pl.S * XB); // partials
  b += Carver::round((size_t)pl.W * XB);        // window sums
  return b + 4096;
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
  MsmPlan pl = msm_plan(n, g2, cfg);
  if (!plan_ok(pl)) {
    set_error("msm: unsupported configuration (window/chunk/segment or size too large)");
    return CSK_ERR_ARG;
  }
  int rc = ctx->ws.ensure(msm_workspace_bytes(pl, g2));
  if (rc) return rc;
  StageTimer *tm = nullptr;
  if (ctx->timing) {
    ctx->timer.reset();
    tm = &ctx->timer;
  }
  ctx->last_plan = pl;
  return g2 ? msm_run_g2(ctx->ws.ptr, pl, d_scalars, d_points, d_out, st, tm)
            : msm_run_g1(ctx->ws.ptr, pl, d_scalars, d_points, d_out, st, tm);
}

}  // namespace csk
