// G1 instantiation of the Pippenger pipeline (separate unit: parallel build).
#include "msm_impl.h"

namespace csk {
int msm_run_g1(void *ws, const MsmPlan &pl, const MsmSeg *segs, hipStream_t st, StageTimer *tm) {
  return msm_run<Fq>(ws, pl, segs, st, tm);
}
int msm_table_build_g1(const uint32_t *pts, uint64_t n, uint64_t n_pad, uint32_t c, uint32_t Wt, uint32_t *table,
                       uint32_t *live, hipStream_t st) {
  return msm_table_build<Fq>(pts, n, n_pad, c, Wt, table, live, st);
}
int msm_run_table_g1(void *ws, const MsmPlan &pl, const MsmSeg *hsegs, const TabArgs &args, uint32_t count,
                     uint32_t max_nt, uint32_t Lrm, uint32_t Sm, hipStream_t st) {
  return msm_run_table<Fq>(ws, pl, hsegs, args, count, max_nt, Lrm, Sm, st);
}
size_t msm_table_ws_bytes(const MsmPlan &pl, bool g2, uint32_t Sm, uint32_t count) {
  return msm_table_workspace_bytes(pl, g2, Sm, count);
}
}  // namespace csk
