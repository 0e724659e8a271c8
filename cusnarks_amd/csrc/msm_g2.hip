// G2 instantiation of the Pippenger pipeline (separate unit: parallel build).
#include "msm_impl.h"

namespace csk {
int msm_run_g2(void *ws, const MsmPlan &pl, const MsmSeg *segs, hipStream_t st, StageTimer *tm) {
  return msm_run<Fq2>(ws, pl, segs, st, tm);
}
int msm_table_build_g2(const uint32_t *pts, uint64_t n, uint64_t n_pad, uint32_t c, uint32_t Wt, uint32_t *table,
                       uint32_t *live, hipStream_t st) {
  return msm_table_build<Fq2>(pts, n, n_pad, c, Wt, table, live, st);
}
int msm_run_table_g2(void *ws, const MsmPlan &pl, const MsmSeg *hsegs, const TabArgs &args, uint32_t count,
                     uint32_t max_nt, uint32_t Lrm, uint32_t Sm, hipStream_t st) {
  return msm_run_table<Fq2>(ws, pl, hsegs, args, count, max_nt, Lrm, Sm, st);
}
}  // namespace csk
