// Pippenger multi-exponentiation: launch plan and entry point.
#pragma once
#include "internal.h"

namespace csk {



MsmPlan msm_plan(uint64_t n, bool g2, const csk_msm_cfg *cfg);
size_t msm_workspace_bytes(const MsmPlan &pl, bool g2);
int msm_run_g1(void *ws, const MsmPlan &pl, const uint32_t *s, const uint32_t *p, uint32_t *out, hipStream_t st, StageTimer *tm);
int msm_run_g2(void *ws, const MsmPlan &pl, const uint32_t *s, const uint32_t *p, uint32_t *out, hipStream_t st, StageTimer *tm);
int msm_launch(csk_ctx *ctx, bool g2, const uint32_t *d_scalars, const uint32_t *d_points, uint64_t n,
               uint32_t *d_out, const csk_msm_cfg *cfg, hipStream_t st);

}  // namespace csk
