// Pippenger multi-exponentiation: launch plan and entry point.
#pragma once
#include "internal.h"

namespace csk {

struct MsmPlan {
  uint32_t c;            // window bits
  uint32_t W;            // windows: floor(254 / c) + 1 signed-digit windows cover k < r < 2^254
  uint32_t B;            // buckets per window, 2^(c-1) (signed digits)
  uint32_t K;            // W * B bucket keys
  uint32_t L;            // sorted entries per accumulation thread
  uint32_t Lr;           // buckets per reduction thread
  uint32_t S;            // reduction segments per window, B / Lr
  uint64_t n;            // points
  uint64_t max_entries;  // n * W (upper bound on nonzero digits)
  uint64_t max_chunks;   // ceil(max_entries / L)
};

MsmPlan msm_plan(uint64_t n, bool g2, const csk_msm_cfg *cfg);
size_t msm_workspace_bytes(const MsmPlan &pl, bool g2);
int msm_run_g1(void *ws, const MsmPlan &pl, const uint32_t *s, const uint32_t *p, uint32_t *out, hipStream_t st);
int msm_run_g2(void *ws, const MsmPlan &pl, const uint32_t *s, const uint32_t *p, uint32_t *out, hipStream_t st);
int msm_launch(csk_ctx *ctx, bool g2, const uint32_t *d_scalars, const uint32_t *d_points, uint64_t n,
               uint32_t *d_out, const csk_msm_cfg *cfg, hipStream_t st);

}  // namespace csk
