// Pippenger multi-exponentiation: launch plans, the per-member descriptor of
// a batched group, and the entry points shared by the single and batched paths.
#pragma once
#include "internal.h"

namespace csk {

// One multiexp inside a group (device-resident table of J+1 entries; entry J
// is a sentinel holding the totals).  All members of a group share c.
struct MsmSeg {
  const uint32_t *scalars;  // n x 8 u32, normal form
  const uint32_t *points;   // n affine points, Montgomery form
  uint32_t *out;            // one affine point
  uint64_t n;               // points of this member
  uint64_t pbase;           // points of the members before it
  uint32_t NT;              // sort tiles per window, ceil(n / 4096)
  uint32_t cbase;           // first pass-1 counter of the member
  uint32_t tbase;           // first pass-1 workgroup of the member
  uint32_t qbase;           // first point the bucket loop reads: pbase, or a table window's slice
};
static_assert(sizeof(MsmSeg) == 56, "MsmSeg layout");

static constexpr uint32_t kSortTile = 4096;  // digits of one window per pass-1 workgroup
// widest window: sort pass 2 keeps 3 * 2^(c-9) words of counters in LDS beside
// its staging (25 KB with the 4096-entry tile, 49 KB with the 8192-entry tile,
// which is used only up to c = 19), so a workgroup stays within 64 KB up to c = 20
static constexpr uint32_t kMaxWindowBits = 20;

// Bytes of the sort buffer: the points in the 29-bit-limb form (64 B G1,
// 128 B G2 each, written with the digits), then the pass-1 arrays (u32 point
// index | sign and u16 low bucket bits per entry), whose space the 29-bit
// bucket and carry slots (144 B G1, 288 B G2 each) reuse once the sort has
// consumed them.
inline uint64_t msm_sortbuf_bytes(const MsmPlan &pl, bool g2) {
  const uint64_t pts = Carver::round(pl.n * (g2 ? 128 : 64));
  const uint64_t sort = Carver::round(pl.max_entries * 4) + pl.max_entries * 2;
  const uint64_t sb = g2 ? 288 : 144;
  const uint64_t slots = Carver::round((uint64_t)pl.J * pl.K * sb) + pl.max_chunks * sb;
  return pts + (sort > slots ? sort : slots);
}

uint32_t msm_choose_c(uint64_t n, bool g2, const csk_ctx *ctx, const csk_msm_cfg *cfg);
// group plan for members of sizes n[0..J) sharing window width c
MsmPlan msm_group_plan(uint32_t c, const uint64_t *n, uint32_t J, bool g2, uint32_t chunk, uint32_t reduce_seg);
bool msm_plan_ok(const MsmPlan &pl);
size_t msm_workspace_bytes(const MsmPlan &pl, bool g2);
// fill the host copy of a group's descriptor table (J+1 entries)
void msm_fill_segs(const MsmPlan &pl, const uint64_t *n, const uint32_t *const *scalars,
                   const uint32_t *const *points, uint32_t *const *out, MsmSeg *segs);
int msm_run_g1(void *ws, const MsmPlan &pl, const MsmSeg *segs, hipStream_t st, StageTimer *tm);
int msm_run_g2(void *ws, const MsmPlan &pl, const MsmSeg *segs, hipStream_t st, StageTimer *tm);
// fixed-base tables (msm_impl.h): build the table of n points for width c
// (Wt windows, n_pad a multiple of the sort tile), and run a table multiexp
int msm_table_build_g1(const uint32_t *pts, uint64_t n, uint64_t n_pad, uint32_t c, uint32_t Wt, uint32_t *table,
                       uint32_t *live, hipStream_t st);
int msm_table_build_g2(const uint32_t *pts, uint64_t n, uint64_t n_pad, uint32_t c, uint32_t Wt, uint32_t *table,
                       uint32_t *live, hipStream_t st);
// Up to kTabBatch table runs in one pipeline, passed by value in the kernel
// arguments: run p multiplies scalar vector s[p] (n[p] scalars) by the points
// of a table whose point-at-infinity bits are live[p], into o[p].  The tables
// share the curve and the window width, so every run has the same Wt.
static constexpr uint32_t kTabBatch = 64;
struct TabArgs {
  const uint32_t *s[kTabBatch];
  uint32_t *o[kTabBatch];
  const uint32_t *live[kTabBatch];
  uint64_t n[kTabBatch];
};
// hsegs: host descriptors of the count * Wt members (m = p * Wt + w: window w
// of run p, n = the table's n_pad, points = that window's slice), copied into
// the workspace; max_nt: the largest n_pad / 4096 among the runs.
int msm_run_table_g1(void *ws, const MsmPlan &pl, const MsmSeg *hsegs, const TabArgs &args, uint32_t count,
                     uint32_t max_nt, uint32_t Lrm, uint32_t Sm, hipStream_t st);
int msm_run_table_g2(void *ws, const MsmPlan &pl, const MsmSeg *hsegs, const TabArgs &args, uint32_t count,
                     uint32_t max_nt, uint32_t Lrm, uint32_t Sm, hipStream_t st);
size_t msm_table_ws_bytes(const MsmPlan &pl, bool g2, uint32_t Sm, uint32_t count);

// looks up this device's copy of the no-allocation scratch array (the device current)
int msm_noalloc_setup(int device);
int msm_launch(csk_ctx *ctx, bool g2, const uint32_t *d_scalars, const uint32_t *d_points, uint64_t n,
               uint32_t *d_out, const csk_msm_cfg *cfg, hipStream_t st);

}  // namespace csk
