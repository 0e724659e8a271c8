// G1 instantiation of the Pippenger pipeline (separate unit: parallel build).
#include "msm_impl.h"

namespace csk {
int msm_run_g1(void *ws, const MsmPlan &pl, const MsmSeg *segs, hipStream_t st, StageTimer *tm) {
  return msm_run<Fq>(ws, pl, segs, st, tm);
}
}  // namespace csk
