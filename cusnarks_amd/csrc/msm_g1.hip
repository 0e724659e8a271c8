// G1 instantiation of the Pippenger pipeline (separate unit: parallel build).
#include "msm_impl.h"

namespace csk {
int msm_run_g1(void *ws, const MsmPlan &pl, const uint32_t *s, const uint32_t *p, uint32_t *out, hipStream_t st, StageTimer *tm) {
  return msm_run<Fq>(ws, pl, s, p, out, st, tm);
}
}  // namespace csk
