// Throughput probes that establish the integer-VALU roofline on this GPU.
//
//   kind 0: independent v_mad_u64_u32 chains (8 per lane) -> MAD64 / s
//   kind 1: independent Fq Montgomery products (4 chains per lane) -> mulmont / s
//   kind 2: the same with the compiler-scheduled CIOS product (comparison)
//   kind 3: a plain device copy of 1 GiB (16-byte vector loads and stores) ->
//           bytes moved (read + written) / s; also the reference kernel for
//           scaling the FETCH_SIZE / WRITE_SIZE counters to bytes
//
// Both run a fixed, in-bounds grid with a compile-time trip count; the result
// is stored so the work cannot be elided.
#include "internal.h"
#include "field.h"
#include "curve.h"

namespace csk {

__global__ void __launch_bounds__(256) k_probe_mad(uint32_t iters, uint64_t *__restrict__ sink) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t a[8];
  for (int j = 0; j < 8; j++) a[j] = 0x9E3779B97F4A7C15ull * (t + 1 + j);
  uint32_t k = 0x9E3779B9u ^ t;
  for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = (uint64_t)(uint32_t)a[j] * k + a[j];
  }
  uint64_t s = 0;
  for (int j = 0; j < 8; j++) s ^= a[j];
  sink[t] = s;
}

template <bool CIOS>
__global__ void __launch_bounds__(256) k_probe_mul(uint32_t iters, uint32_t *__restrict__ sink) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  Fq x[4];
  for (int j = 0; j < 4; j++)
    for (int i = 0; i < 8; i++) x[j].v[i] = (0x9E3779B9u * (t + 7 * j + i + 1)) >> 3;
  Fq y = x[3];
  for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
    for (int j = 0; j < 4; j++) x[j] = CIOS ? Fq::mul_cios(x[j], y) : Fq::mul(x[j], y);
  }
  Fq s = Fq::add(Fq::add(x[0], x[1]), Fq::add(x[2], x[3]));
  st8(sink + 8 * (uint64_t)t, s.v);
}

__global__ void __launch_bounds__(256) k_probe_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                    uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace csk

using namespace csk;

extern "C" int csk_valu_probe(csk_ctx *ctx, uint32_t kind, double *ops_per_s, double *ms) {
  if (!ctx || !ops_per_s || kind > 3) {
    set_error("valu_probe: bad argument");
    return CSK_ERR_ARG;
  }
  CSK_HIP(hipSetDevice(ctx->device));
  hipDeviceProp_t prop;
  CSK_HIP(hipGetDeviceProperties(&prop, ctx->device));
  const uint32_t blocks = (uint32_t)prop.multiProcessorCount * 32;  // 32 blocks of 256 per CU
  const uint32_t threads = 256;
  const uint32_t iters = kind == 0 ? 4096 : 64;
  const uint64_t copy_bytes = 1ull << 30;
  int rc = ctx->ws.ensure(kind == 3 ? 2 * copy_bytes : (size_t)blocks * threads * 32);
  if (rc) return rc;
  hipEvent_t e0, e1;
  CSK_HIP(hipEventCreate(&e0));
  CSK_HIP(hipEventCreate(&e1));
  double best = 1e30;
  for (int rep = 0; rep < 3; rep++) {
    CSK_HIP(hipEventRecord(e0, nullptr));
    if (kind == 0) k_probe_mad<<<blocks, threads>>>(iters, (uint64_t *)ctx->ws.ptr);
    else if (kind == 1) k_probe_mul<false><<<blocks, threads>>>(iters, (uint32_t *)ctx->ws.ptr);
    else if (kind == 2) k_probe_mul<true><<<blocks, threads>>>(iters, (uint32_t *)ctx->ws.ptr);
    else
      k_probe_copy<<<blocks, threads>>>((const uint4 *)ctx->ws.ptr, (uint4 *)((char *)ctx->ws.ptr + copy_bytes),
                                        copy_bytes / 16);
    CSK_CHECK_LAUNCH();
    CSK_HIP(hipEventRecord(e1, nullptr));
    CSK_HIP(hipEventSynchronize(e1));
    float m = 0;
    CSK_HIP(hipEventElapsedTime(&m, e0, e1));
    if (m < best) best = m;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  double ops = kind == 3 ? 2.0 * copy_bytes : (double)blocks * threads * iters * (kind == 0 ? 8 : 4);
  *ops_per_s = ops / (best * 1e-3);
  if (ms) *ms = best;
  return CSK_OK;
}
