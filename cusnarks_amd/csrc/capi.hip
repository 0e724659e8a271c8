// C ABI of libcsk.so: contexts, memory helpers, elementwise field kernels,
// Jacobian->affine, fixed-base generation and the multiexp entry points.
#include <cstring>
#include <string>
#include "internal.h"
#include "curve.h"
#include "msm.h"

namespace csk {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

int Workspace::ensure(size_t need) {
  if (need <= bytes) return CSK_OK;
  if (ptr) {
    // work on any stream may still use the old buffer
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    if (e != hipSuccess) {
      set_error(std::string("workspace free: ") + hipGetErrorString(e));
      return CSK_ERR_HIP;
    }
  }
  size_t want = need + need / 8;
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    set_error("workspace alloc of " + std::to_string(want) + " bytes: " + hipGetErrorString(e));
    return CSK_ERR_NOMEM;
  }
  bytes = want;
  gen++;
  return CSK_OK;
}

void Workspace::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

// ------------------------------------------------------------------ field ops
template <int PIDX>
__global__ void __launch_bounds__(256) k_field_op(int op, const uint32_t *__restrict__ a,
                                                  const uint32_t *__restrict__ b, uint32_t *__restrict__ out,
                                                  uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  using F = Fp<PIDX>;
  F x, y, r;
  ld8(x.v, a + 8 * i);
  if (op <= CSK_OP_SUBM) ld8(y.v, b + 8 * i);
  switch (op) {
    case CSK_OP_MULM: r = F::mul(x, y); break;
    case CSK_OP_ADDM: r = F::add(x, y); break;
    case CSK_OP_SUBM: r = F::sub(x, y); break;
    case CSK_OP_TOMONT: r = F::to_mont(x); break;
    case CSK_OP_FROMMONT: r = F::from_mont(x); break;
    default: r = F::inv(x); break;
  }
  st8(out + 8 * i, r.v);
}

__global__ void __launch_bounds__(256) k_mul_ext(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                                                 uint32_t *__restrict__ out, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fq2 x = FieldIO<Fq2>::load(a + 16 * i), y = FieldIO<Fq2>::load(b + 16 * i);
  FieldIO<Fq2>::store(out + 16 * i, Fq2::mul(x, y));
}

// ------------------------------------------------------------------ jac2aff
template <class F>
__global__ void __launch_bounds__(256) k_jac2aff(const uint32_t *__restrict__ in, uint64_t n,
                                                 uint32_t *__restrict__ out, uint32_t strip_last) {
  constexpr int NW = FieldIO<F>::NW;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t *p = in + i * 3 * NW;
  uint32_t *o = out + i * (strip_last ? 2 : 3) * NW;
  F X = FieldIO<F>::load(p), Y = FieldIO<F>::load(p + NW), Z = FieldIO<F>::load(p + 2 * NW);
  F x, y;
  if (Z.is_zero()) {
    x = F::zero();
    y = F::one();
  } else {
    F zi = F::inv(Z);
    F zi2 = F::sqr(zi);
    x = F::mul(X, zi2);
    y = F::mul(Y, F::mul(zi2, zi));
  }
  FieldIO<F>::store(o, x);
  FieldIO<F>::store(o + NW, y);
  if (!strip_last) FieldIO<F>::store(o + 2 * NW, Z.is_zero() ? F::zero() : F::one());
}

// ------------------------------------------------------------------ fixed base
// table[w][d] = d * 2^(8w) * G (affine), w < 32, 1 <= d < 256 (entry 0 unused)
template <class F> __device__ Affine<F> generator();
template <> __device__ Affine<Fq> generator<Fq>() {
  Affine<Fq> g;
  g.x = Fq::one();
  const uint32_t two[8] = {2334006074u, 2797242139u, 3951957627u, 351393361u,
                           4042427480u, 3437053662u, 873447006u,  471134083u};
  for (int i = 0; i < 8; i++) g.y.v[i] = two[i];
  return g;
}
template <> __device__ Affine<Fq2> generator<Fq2>() {
  const uint32_t v[4][8] = {
      {45883430u, 2390996433u, 1232798066u, 3706394933u, 2541820639u, 4223149639u, 2945863739u, 425146433u},
      {2823577920u, 2947838845u, 1476581572u, 1615060314u, 1386229638u, 166285564u, 988445547u, 352252035u},
      {2288773622u, 1637743261u, 4120812408u, 4269789847u, 589004286u, 4288551522u, 2929607174u, 687701739u},
      {3340261102u, 1678334806u, 847068347u, 3696752930u, 859115638u, 1442395582u, 2482857090u, 228892902u}};
  Affine<Fq2> g;
  for (int i = 0; i < 8; i++) {
    g.x.c0.v[i] = v[0][i];
    g.x.c1.v[i] = v[1][i];
    g.y.c0.v[i] = v[2][i];
    g.y.c1.v[i] = v[3][i];
  }
  return g;
}

template <class F> __global__ void k_fb_table_xyzz(uint32_t *__restrict__ tmp) {
  constexpr int XW = 4 * FieldIO<F>::NW;
  uint32_t w = threadIdx.x;
  if (w >= 32) return;
  XYZZ<F> base = XYZZ<F>::from_affine(generator<F>());
  for (uint32_t j = 0; j < 8 * w; j++) base = xyzz_dbl(base);
  Affine<F> ba = xyzz_to_affine(base);
  XYZZ<F> acc = XYZZ<F>::inf();
  for (uint32_t d = 1; d < 256; d++) {
    acc = xyzz_madd(acc, ba);
    store_xyzz(tmp + ((uint64_t)w * 256 + d) * XW, acc);
  }
}

template <class F> __global__ void __launch_bounds__(256) k_fb_table_aff(const uint32_t *__restrict__ tmp,
                                                                         uint32_t *__restrict__ table) {
  constexpr int XW = 4 * FieldIO<F>::NW;
  constexpr int PW = 2 * FieldIO<F>::NW;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 32 * 256 || (i & 255) == 0) return;
  store_affine(table + (uint64_t)i * PW, xyzz_to_affine(load_xyzz<F>(tmp + (uint64_t)i * XW)));
}

template <class F>
__global__ void __launch_bounds__(256) k_fixed_base(const uint32_t *__restrict__ table,
                                                    const uint32_t *__restrict__ scalars, uint64_t n,
                                                    uint32_t *__restrict__ out) {
  constexpr int PW = 2 * FieldIO<F>::NW;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t k[8];
  ld8(k, scalars + 8 * i);
  XYZZ<F> acc = XYZZ<F>::inf();
  for (uint32_t w = 0; w < 32; w++) {
    uint32_t d = (k[w >> 2] >> (8 * (w & 3))) & 255u;
    if (d) acc = xyzz_madd(acc, load_affine<F>(table + ((uint64_t)w * 256 + d) * PW));
  }
  store_affine(out + i * PW, xyzz_to_affine(acc));
}

template <class F> static int fixed_base(csk_ctx *ctx, uint32_t **tbl, const uint32_t *d_scalars, uint64_t n,
                                         uint32_t *d_out, hipStream_t st) {
  constexpr int XW = 4 * FieldIO<F>::NW;
  constexpr int PW = 2 * FieldIO<F>::NW;
  CSK_HIP(hipSetDevice(ctx->device));
  if (!*tbl) {
    uint32_t *tmp = nullptr, *t = nullptr;
    CSK_HIP(hipMalloc(&tmp, (size_t)32 * 256 * XW * 4));
    hipError_t e = hipMalloc(&t, (size_t)32 * 256 * PW * 4);
    if (e != hipSuccess) {
      (void)hipFree(tmp);
      set_error(std::string("fixed-base table: ") + hipGetErrorString(e));
      return CSK_ERR_NOMEM;
    }
    k_fb_table_xyzz<F><<<1, 64, 0, st>>>(tmp);
    k_fb_table_aff<F><<<32, 256, 0, st>>>(tmp, t);
    CSK_CHECK_LAUNCH();
    CSK_HIP(hipStreamSynchronize(st));
    CSK_HIP(hipFree(tmp));
    *tbl = t;
  }
  if (n == 0) return CSK_OK;
  k_fixed_base<F><<<(uint32_t)((n + 255) / 256), 256, 0, st>>>(*tbl, d_scalars, n, d_out);
  CSK_CHECK_LAUNCH();
  return CSK_OK;
}

// ------------------------------------------------------------------ point checks
// Curve constant b in Montgomery form: 3 for G1, 3 / (9 + u) for the G2 twist.
template <class F> __device__ F curve_b();
template <> __device__ Fq curve_b<Fq>() {
  const uint32_t b[8] = {1353525463u, 2048379561u, 3780452793u, 527090042u,
                         1768673924u, 860613198u,  3457654158u, 706701124u};
  Fq r;
  for (int i = 0; i < 8; i++) r.v[i] = b[i];
  return r;
}
template <> __device__ Fq2 curve_b<Fq2>() {
  const uint32_t b[2][8] = {
      {2008548008u, 1006188771u, 909333341u, 34282279u, 1232425568u, 649588208u, 1132767341u, 622118450u},
      {3520921447u, 954723532u, 2479754558u, 1710273405u, 581697706u, 3611939037u, 1248365901u, 21084622u}};
  Fq2 r;
  for (int i = 0; i < 8; i++) {
    r.c0.v[i] = b[0][i];
    r.c1.v[i] = b[1][i];
  }
  return r;
}

__device__ __forceinline__ bool canonical_q(const uint32_t *v) {
  for (int i = 7; i >= 0; i--)
    if (v[i] != FieldConst<0>::p[i]) return v[i] < FieldConst<0>::p[i];
  return false;
}
__device__ __forceinline__ bool canonical(const Fq &a) { return canonical_q(a.v); }
__device__ __forceinline__ bool canonical(const Fq2 &a) { return canonical_q(a.c0.v) && canonical_q(a.c1.v); }

// res[0] += number of bad points, res[1] = min index of a bad point.  A point
// is bad if a coordinate is not below q, if it is not on the curve, or (with
// subgroup set) if [r]P is not the identity.  Both infinity encodings pass.
template <class F>
__global__ void __launch_bounds__(256) k_ec_check(const uint32_t *__restrict__ pts, uint64_t n, uint32_t subgroup,
                                                  unsigned long long *__restrict__ res) {
  constexpr int PW = 2 * FieldIO<F>::NW;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Affine<F> p = load_affine<F>(pts + i * PW);
  if (affine_is_inf(p)) return;
  bool bad = !canonical(p.x) || !canonical(p.y);
  if (!bad) {
    F rhs = F::add(F::mul(F::sqr(p.x), p.x), curve_b<F>());
    bad = !(F::sqr(p.y) == rhs);
  }
  if (!bad && subgroup) {
    XYZZ<F> acc = XYZZ<F>::inf();
    for (int b = 253; b >= 0; b--) {  // r < 2^254, top bit 253
      acc = xyzz_dbl(acc);
      if ((FieldConst<1>::p[b >> 5] >> (b & 31)) & 1u) acc = xyzz_madd(acc, p);
    }
    bad = !acc.is_inf();
  }
  if (bad) {
    atomicAdd(&res[0], 1ull);
    atomicMin(&res[1], (unsigned long long)i);
  }
}

__global__ void k_ec_check_init(unsigned long long *res) {
  if (threadIdx.x == 0) {
    res[0] = 0;
    res[1] = ~0ull;
  }
}

}  // namespace csk

using namespace csk;

extern "C" {

const char *csk_last_error(void) { return g_last_error.c_str(); }

int csk_device_count(int *count) {
  if (!count) return CSK_ERR_ARG;
  CSK_HIP(hipGetDeviceCount(count));
  return CSK_OK;
}

int csk_ctx_create(int device, csk_ctx **out) {
  if (!out) {
    set_error("ctx_create: null out");
    return CSK_ERR_ARG;
  }
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    set_error("no HIP device available");
    return CSK_ERR_NODEV;
  }
  if (device < 0 || device >= n) {
    set_error("device index out of range");
    return CSK_ERR_ARG;
  }
  CSK_HIP(hipSetDevice(device));
  // the no-allocation multiexp's scratch address on this device; on failure
  // that entry point retries on its first call and reports the error there
  (void)msm_noalloc_setup(device);
  csk_ctx *c = new csk_ctx();
  c->device = device;
  *out = c;
  return CSK_OK;
}

void csk_ctx_destroy(csk_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  ctx->ws.release();
  ctx->ws2.release();
  ctx->batch_ws.release();
  if (ctx->g1_table) (void)hipFree(ctx->g1_table);
  if (ctx->g2_table) (void)hipFree(ctx->g2_table);
  if (ctx->check_res) (void)hipFree(ctx->check_res);
  ctx->timer.destroy();
  delete ctx;
}

int csk_ctx_set_timing(csk_ctx *ctx, int on) {
  if (!ctx) return CSK_ERR_ARG;
  ctx->timing = on != 0;
  return CSK_OK;
}

int csk_last_stage_times(csk_ctx *ctx, float *ms, int n) {
  if (!ctx || !ms || n < 1) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  float v[4] = {ctx->timer.between(0, 1), ctx->timer.between(1, 2), ctx->timer.between(2, 3),
                ctx->timer.between(0, 3)};
  for (int i = 0; i < n && i < 4; i++) ms[i] = v[i];
  return CSK_OK;
}

int csk_last_msm_plan(csk_ctx *ctx, uint32_t *out, int n) {
  if (!ctx || !out) return CSK_ERR_ARG;
  const csk::MsmPlan &p = ctx->last_plan;
  uint64_t v[8] = {p.c, p.W, p.B, p.L, p.Lr, p.S, p.n, p.max_entries};
  for (int i = 0; i < n && i < 8; i++) out[i] = (uint32_t)v[i];
  return CSK_OK;
}

int csk_malloc(csk_ctx *ctx, void **d_ptr, uint64_t bytes) {
  if (!ctx || !d_ptr) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    set_error(std::string("csk_malloc: ") + hipGetErrorString(e));
    return CSK_ERR_NOMEM;
  }
  return CSK_OK;
}

int csk_free(csk_ctx *ctx, void *d_ptr) {
  if (!ctx) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  CSK_HIP(hipFree(d_ptr));
  return CSK_OK;
}

int csk_host_alloc(csk_ctx *ctx, void **h_ptr, uint64_t bytes) {
  if (!ctx || !h_ptr) return CSK_ERR_ARG;
  *h_ptr = nullptr;
  CSK_HIP(hipSetDevice(ctx->device));
  hipError_t e = hipHostMalloc(h_ptr, bytes ? bytes : 1, hipHostMallocDefault);
  if (e != hipSuccess) {
    *h_ptr = nullptr;
    set_error(std::string("csk_host_alloc: ") + hipGetErrorString(e));
    return CSK_ERR_NOMEM;
  }
  return CSK_OK;
}

int csk_host_free(csk_ctx *ctx, void *h_ptr) {
  if (!ctx) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  CSK_HIP(hipHostFree(h_ptr));
  return CSK_OK;
}

int csk_memcpy_h2d(csk_ctx *ctx, void *d_dst, const void *src, uint64_t bytes, void *stream) {
  if (!ctx) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  CSK_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  return CSK_OK;
}

int csk_memcpy_d2h(csk_ctx *ctx, void *dst, const void *d_src, uint64_t bytes, void *stream) {
  if (!ctx) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  CSK_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  CSK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return CSK_OK;
}

int csk_memcpy_d2d(csk_ctx *ctx, void *d_dst, const void *d_src, uint64_t bytes, void *stream) {
  if (!ctx) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  CSK_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return CSK_OK;
}

int csk_stream_sync(csk_ctx *ctx, void *stream) {
  if (!ctx) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  CSK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return CSK_OK;
}

int csk_stream_create(csk_ctx *ctx, void **stream) {
  if (!ctx || !stream) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  hipStream_t s = nullptr;
  CSK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return CSK_OK;
}

int csk_stream_destroy(csk_ctx *ctx, void *stream) {
  if (!ctx || !stream) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  CSK_HIP(hipStreamDestroy((hipStream_t)stream));
  return CSK_OK;
}

int csk_stream_wait_event(csk_ctx *ctx, void *stream, void *event) {
  if (!ctx || !event) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  CSK_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
  return CSK_OK;
}

int csk_event_create(csk_ctx *ctx, void **event) {
  if (!ctx || !event) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  hipEvent_t e = nullptr;
  CSK_HIP(hipEventCreate(&e));
  *event = e;
  return CSK_OK;
}

int csk_event_destroy(csk_ctx *ctx, void *event) {
  if (!ctx || !event) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  CSK_HIP(hipEventDestroy((hipEvent_t)event));
  return CSK_OK;
}

int csk_event_record(csk_ctx *ctx, void *event, void *stream) {
  if (!ctx || !event) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  CSK_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return CSK_OK;
}

int csk_event_query(csk_ctx *ctx, void *event) {
  if (!ctx || !event) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  hipError_t e = hipEventQuery((hipEvent_t)event);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();  // not-ready is a status, not a sticky error
    return 0;
  }
  set_error(std::string("event query: ") + hipGetErrorString(e));
  return CSK_ERR_HIP;
}

int csk_event_sync(csk_ctx *ctx, void *event) {
  if (!ctx || !event) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  CSK_HIP(hipEventSynchronize((hipEvent_t)event));
  return CSK_OK;
}

int csk_event_elapsed(csk_ctx *ctx, void *start, void *end, float *ms) {
  if (!ctx || !start || !end || !ms) return CSK_ERR_ARG;
  CSK_HIP(hipSetDevice(ctx->device));
  CSK_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
  return CSK_OK;
}

int csk_msm_g1(csk_ctx *ctx, const uint32_t *d_scalars, const uint32_t *d_points, uint64_t n, uint32_t *d_out,
               const csk_msm_cfg *cfg, void *stream) {
  if (!ctx) return CSK_ERR_ARG;
  return msm_launch(ctx, false, d_scalars, d_points, n, d_out, cfg, (hipStream_t)stream);
}

int csk_msm_g2(csk_ctx *ctx, const uint32_t *d_scalars, const uint32_t *d_points, uint64_t n, uint32_t *d_out,
               const csk_msm_cfg *cfg, void *stream) {
  if (!ctx) return CSK_ERR_ARG;
  return msm_launch(ctx, true, d_scalars, d_points, n, d_out, cfg, (hipStream_t)stream);
}

static int msm_host(csk_ctx *ctx, bool g2, const uint32_t *scalars, const uint32_t *points, uint64_t n,
                    uint32_t *out, const csk_msm_cfg *cfg) {
  if (!ctx || !out || (n && (!scalars || !points))) {
    set_error("msm_host: null argument");
    return CSK_ERR_ARG;
  }
  CSK_HIP(hipSetDevice(ctx->device));
  const size_t pw = g2 ? 32 : 16;
  uint32_t *d = nullptr;
  size_t bytes = n * 8 * 4 + n * pw * 4 + pw * 4;
  hipError_t e = hipMalloc(&d, bytes);
  if (e != hipSuccess) {
    set_error(std::string("msm_host alloc: ") + hipGetErrorString(e));
    return CSK_ERR_NOMEM;
  }
  uint32_t *ds = d, *dp = d + n * 8, *dout = d + n * 8 + n * pw;
  int rc = CSK_OK;
  if (n) {
    e = hipMemcpy(ds, scalars, n * 32, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp, points, n * pw * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      set_error(std::string("msm_host copy: ") + hipGetErrorString(e));
      rc = CSK_ERR_HIP;
    }
  }
  if (rc == CSK_OK) rc = msm_launch(ctx, g2, ds, dp, n, dout, cfg, nullptr);
  if (rc == CSK_OK) {
    e = hipMemcpy(out, dout, pw * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      set_error(std::string("msm_host copy back: ") + hipGetErrorString(e));
      rc = CSK_ERR_HIP;
    }
  }
  (void)hipFree(d);
  return rc;
}

int csk_msm_g1_host(csk_ctx *ctx, const uint32_t *scalars, const uint32_t *points, uint64_t n, uint32_t *out,
                    const csk_msm_cfg *cfg) {
  return msm_host(ctx, false, scalars, points, n, out, cfg);
}

int csk_msm_g2_host(csk_ctx *ctx, const uint32_t *scalars, const uint32_t *points, uint64_t n, uint32_t *out,
                    const csk_msm_cfg *cfg) {
  return msm_host(ctx, true, scalars, points, n, out, cfg);
}

int csk_field_op(csk_ctx *ctx, int op, uint32_t pidx, const uint32_t *d_a, const uint32_t *d_b, uint32_t *d_out,
                 uint64_t n, void *stream) {
  if (!ctx || op < 0 || op > CSK_OP_MULM_EXT || pidx > 1 || (n && (!d_a || !d_out))) {
    set_error("field_op: bad argument");
    return CSK_ERR_ARG;
  }
  if ((op <= CSK_OP_SUBM || op == CSK_OP_MULM_EXT) && n && !d_b) {
    set_error("field_op: missing second operand");
    return CSK_ERR_ARG;
  }
  if (op == CSK_OP_MULM_EXT && pidx != 0) {
    set_error("field_op: Fq2 product is defined over the base field (pidx 0) only");
    return CSK_ERR_ARG;
  }
  if (n == 0) return CSK_OK;
  CSK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  uint32_t g = (uint32_t)((n + 255) / 256);
  if (op == CSK_OP_MULM_EXT) k_mul_ext<<<g, 256, 0, st>>>(d_a, d_b, d_out, n);
  else if (pidx == 0) k_field_op<0><<<g, 256, 0, st>>>(op, d_a, d_b, d_out, n);
  else k_field_op<1><<<g, 256, 0, st>>>(op, d_a, d_b, d_out, n);
  CSK_CHECK_LAUNCH();
  return CSK_OK;
}

int csk_ec_jac2aff(csk_ctx *ctx, uint32_t ec2, const uint32_t *d_jac, uint64_t n, uint32_t *d_out,
                   uint32_t strip_last, void *stream) {
  if (!ctx || (n && (!d_jac || !d_out))) return CSK_ERR_ARG;
  if (n == 0) return CSK_OK;
  CSK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  uint32_t g = (uint32_t)((n + 255) / 256);
  if (ec2) k_jac2aff<Fq2><<<g, 256, 0, st>>>(d_jac, n, d_out, strip_last);
  else k_jac2aff<Fq><<<g, 256, 0, st>>>(d_jac, n, d_out, strip_last);
  CSK_CHECK_LAUNCH();
  return CSK_OK;
}

int csk_fixed_base_g1(csk_ctx *ctx, const uint32_t *d_scalars, uint64_t n, uint32_t *d_out, void *stream) {
  if (!ctx || (n && (!d_scalars || !d_out))) return CSK_ERR_ARG;
  return fixed_base<Fq>(ctx, &ctx->g1_table, d_scalars, n, d_out, (hipStream_t)stream);
}

int csk_fixed_base_g2(csk_ctx *ctx, const uint32_t *d_scalars, uint64_t n, uint32_t *d_out, void *stream) {
  if (!ctx || (n && (!d_scalars || !d_out))) return CSK_ERR_ARG;
  return fixed_base<Fq2>(ctx, &ctx->g2_table, d_scalars, n, d_out, (hipStream_t)stream);
}

int csk_ec_check(csk_ctx *ctx, uint32_t ec2, const uint32_t *d_points, uint64_t n, uint32_t subgroup,
                 uint64_t *out, void *stream) {
  if (!ctx || !out || (n && !d_points)) return CSK_ERR_ARG;
  out[0] = 0;
  out[1] = ~0ull;
  if (n == 0) return CSK_OK;
  CSK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (!ctx->check_res) CSK_HIP(hipMalloc(&ctx->check_res, 2 * sizeof(unsigned long long)));
  k_ec_check_init<<<1, 64, 0, st>>>(ctx->check_res);
  uint32_t g = (uint32_t)((n + 255) / 256);
  if (ec2) k_ec_check<Fq2><<<g, 256, 0, st>>>(d_points, n, subgroup, ctx->check_res);
  else k_ec_check<Fq><<<g, 256, 0, st>>>(d_points, n, subgroup, ctx->check_res);
  CSK_CHECK_LAUNCH();
  unsigned long long h[2];
  CSK_HIP(hipMemcpyAsync(h, ctx->check_res, sizeof(h), hipMemcpyDeviceToHost, st));
  CSK_HIP(hipStreamSynchronize(st));
  out[0] = h[0];
  out[1] = h[1];
  return CSK_OK;
}

}  // extern "C"
