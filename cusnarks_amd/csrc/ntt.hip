// Fr number-theoretic transforms, the QAP H step and QAP evaluation on gfx950.
//
// NTT:  X[k] = sum_i a[i] w^(ik), w = 5^((r-1)/N) (cusnarks proots_field,
//       src/cuda/constants.cpp:323-352), natural order in and out, Montgomery
//       data -- the value of ntt_h (src/cuda/ntt.cpp:453-463).
// INTT: N^-1 sum_k X[k] w^(-ik), output Montgomery (fmt 1) or normal (fmt 0),
//       the value of intt_h (ntt.cpp:583-603, IScalerMont/IScalerExt).
//
// Schedule: radix-2 decimation-in-frequency stages L-1 .. 0 grouped into
// passes of up to 11 stages; each pass loads 2^R-point sub-transforms (times G
// adjacent columns, 2048 elements = 64 KB) into LDS, runs its stages there
// with twiddles w_N^k from a device table, and writes back in place.  A
// finalize kernel applies the bit reversal (and for the inverse the index
// reversal i -> -i, since INTT(X)[i] = N^-1 NTT(X)[-i]), the N^-1 scaling and
// optional coset factors in one gather.
#include "internal.h"
#include "field.h"
#include "curve.h"

namespace csk {

static constexpr int NTT_TILE = 2048;  // elements per workgroup pass
static constexpr int NTT_THREADS = 256;
static constexpr uint32_t NTT_MAX_LG = 27;

__constant__ uint32_t kRoot28[8] = {2161196444u, 1668182869u, 608567254u, 2720789364u,
                                    514982872u,  1447373504u, 694417895u, 409005972u};

// rt[j] = w_(2^j) for j = 0..28 (rt[0] = 1)
__global__ void k_roots(uint32_t *__restrict__ rt) {
  if (threadIdx.x || blockIdx.x) return;
  Fr w;
  for (int i = 0; i < 8; i++) w.v[i] = kRoot28[i];
  for (int j = 28; j >= 0; j--) {
    st8(rt + 8 * j, w.v);
    w = Fr::sqr(w);
  }
}

// tw[k] = w_N^k for k < N/2, N = 2^L (w_N^(2^b) = w_(2^(L-b)))
__global__ void __launch_bounds__(256) k_twiddles(const uint32_t *__restrict__ rt, uint32_t L,
                                                  uint32_t *__restrict__ tw) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (1u << (L - 1))) return;
  Fr acc = Fr::one();
  for (uint32_t b = 0; b + 1 < L; b++) {
    if ((k >> b) & 1u) {
      Fr f;
      ld8(f.v, rt + 8 * (L - b));
      acc = Fr::mul(acc, f);
    }
  }
  st8(tw + 8 * (uint64_t)k, acc.v);
}

// One DIF pass: stages s_hi .. s_hi-R+1.  Sub-problem sp (bits of the index
// outside [s_lo, s_hi]) holds elements sp_low + (t << s_lo) + sp_high, t < 2^R.
// A workgroup takes G = TILE / 2^R consecutive sub-problems.  tw is the table
// for N_tab = 2^Ltab >= N (stride 2^(Ltab-L)).  TILE elements (32 B each) sit
// in LDS: a 2048-element tile takes 64 KB (two workgroups per CU), a 1024-
// element one 32 KB (five per CU, more waves to hide the LDS and twiddle
// latency of each butterfly).
template <int TILE>
__global__ void __launch_bounds__(NTT_THREADS) k_ntt_dif_pass(const uint32_t *__restrict__ src,
                                                              uint32_t *__restrict__ dst, uint32_t L,
                                                              uint32_t s_hi, uint32_t R,
                                                              const uint32_t *__restrict__ tw,
                                                              uint32_t Ltab) {
  __shared__ uint32_t sh[TILE * 8];
  const uint32_t s_lo = s_hi + 1 - R;
  const uint32_t sub = 1u << R;
  const uint32_t G = TILE >> R;  // columns (sub-problems) per workgroup
  const uint32_t lowmask = (1u << s_lo) - 1;
  const uint32_t sp0 = blockIdx.x * G;
  // element address of (g, t)
  auto addr = [&](uint32_t g, uint32_t t) -> uint64_t {
    uint32_t sp = sp0 + g;
    uint64_t low = sp & lowmask, high = (uint64_t)(sp >> s_lo) << (s_hi + 1);
    return high | ((uint64_t)t << s_lo) | low;
  };
  // load: consecutive threads take consecutive addresses
  const bool col_major = s_lo > 0;  // G adjacent columns share each t
  for (uint32_t i = threadIdx.x; i < TILE; i += NTT_THREADS) {
    uint32_t g, t;
    if (col_major) { g = i % G; t = i / G; }
    else { g = i / sub; t = i % sub; }
    uint4 *d = reinterpret_cast<uint4 *>(sh + 8 * (t * G + g));
    const uint4 *s = reinterpret_cast<const uint4 *>(src + 8 * addr(g, t));
    d[0] = s[0];
    d[1] = s[1];
  }
  __syncthreads();
  const uint32_t tshift = Ltab - L;
  for (int s = (int)s_hi; s >= (int)s_lo; s--) {
    const uint32_t half = 1u << (s - s_lo);
    for (uint32_t b = threadIdx.x; b < TILE / 2; b += NTT_THREADS) {
      uint32_t g = b % G, q = b / G;
      uint32_t t = ((q / half) * 2 * half) + (q % half);  // bit (s - s_lo) of t is 0
      uint32_t low = (sp0 + g) & lowmask;
      uint64_t jm = (uint64_t)low + ((uint64_t)(t & (half - 1)) << s_lo);  // j mod 2^s
      uint64_t ti = (jm << (L - 1 - s)) << tshift;
      Fr a, c, w;
      uint32_t *pa = sh + 8 * (t * G + g), *pc = sh + 8 * ((t + half) * G + g);
      for (int k = 0; k < 8; k++) { a.v[k] = pa[k]; c.v[k] = pc[k]; }
      ld8(w.v, tw + 8 * ti);
      Fr u = Fr::add(a, c);
      Fr v = Fr::mul(Fr::sub(a, c), w);
      for (int k = 0; k < 8; k++) { pa[k] = u.v[k]; pc[k] = v.v[k]; }
    }
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < TILE; i += NTT_THREADS) {
    uint32_t g, t;
    if (col_major) { g = i % G; t = i / G; }
    else { g = i / sub; t = i % sub; }
    const uint4 *s = reinterpret_cast<const uint4 *>(sh + 8 * (t * G + g));
    uint4 *d = reinterpret_cast<uint4 *>(dst + 8 * addr(g, t));
    d[0] = s[0];
    d[1] = s[1];
  }
}

__device__ __forceinline__ uint32_t bitrev(uint32_t x, uint32_t L) { return L ? __brev(x) >> (32 - L) : 0u; }

// dst[i] = scale * src[bitrev(inverse ? -i : i)] * (coset ? w_tab^(i << cshift) : 1)
// (+ optional pointwise product with mul_by[i] before writing at dst[i*ostride+ooff])
__global__ void __launch_bounds__(256) k_ntt_finalize(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                                      uint32_t L, uint32_t inverse,
                                                      const uint32_t *__restrict__ scale,
                                                      const uint32_t *__restrict__ coset_tw, uint32_t cshift,
                                                      const uint32_t *__restrict__ mul_by, uint32_t ostride,
                                                      uint32_t ooff) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t N = 1u << L;
  if (i >= N) return;
  uint32_t j = inverse ? ((N - i) & (N - 1)) : i;
  // every load is issued before the first product, so they are in flight together
  Fr x, s = {}, w = {}, m = {};
  ld8(x.v, src + 8 * (uint64_t)bitrev(j, L));
  if (scale) ld8(s.v, scale);
  if (coset_tw) ld8(w.v, coset_tw + 8 * ((uint64_t)i << cshift));
  if (mul_by) ld8(m.v, mul_by + 8 * (uint64_t)i);
  if (scale) x = Fr::mul(x, s);
  if (coset_tw) x = Fr::mul(x, w);
  if (mul_by) x = Fr::mul(x, m);
  st8(dst + 8 * ((uint64_t)i * ostride + ooff), x.v);
}

// out[i*ostride + ooff] = a[i] * b[i]
__global__ void __launch_bounds__(256) k_pointwise(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                                                   uint32_t *__restrict__ out, uint64_t n, uint32_t ostride,
                                                   uint32_t ooff) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fr x, y;
  ld8(x.v, a + 8 * i);
  ld8(y.v, b + 8 * i);
  st8(out + 8 * (i * ostride + ooff), Fr::mul(x, y).v);
}

// dst[0..nwords) = 0 (a kernel rather than a memset node, so captured graphs replay it correctly)
__global__ void __launch_bounds__(256) k_fill0(uint32_t *__restrict__ dst, uint64_t nwords) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nwords) dst[i] = 0;
}

// scalers: [0] N^-1 R (Montgomery out), [1] N^-1 (normal out)
__global__ void k_scalers(uint32_t L, uint32_t *__restrict__ out) {
  if (threadIdx.x || blockIdx.x) return;
  Fr n = Fr::zero();
  n.v[L >> 5] = 1u << (L & 31);
  Fr inv = Fr::inv(Fr::to_mont(n));  // N^-1 R
  st8(out, inv.v);
  st8(out + 8, Fr::from_mont(inv).v);
}

// ------------------------------------------------------------------ QAP evaluation
// out[row] = R * sum_e val[e] * w[col[e]] over the row's entries (val Montgomery,
// w normal): one Montgomery product per entry gives the normal-form term, the
// row sum is moved to Montgomery form once.  Same value as mpoly_eval_h
// (cusnarks src/cuda/mpoly.cpp:125-168): sum coeff * toMont(w).
__global__ void __launch_bounds__(256) k_spmv(const uint64_t *__restrict__ rowptr, const uint32_t *__restrict__ col,
                                              const uint32_t *__restrict__ val, const uint32_t *__restrict__ wit,
                                              uint32_t nrows, uint32_t *__restrict__ out) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  Fr acc = Fr::zero();
  for (uint64_t e = rowptr[r]; e < rowptr[r + 1]; e++) {
    Fr v, w;
    ld8(v.v, val + 8 * e);
    ld8(w.v, wit + 8 * (uint64_t)col[e]);
    acc = Fr::add(acc, Fr::mul(v, w));
  }
  st8(out + 8 * (uint64_t)r, Fr::to_mont(acc).v);
}

// ------------------------------------------------------------------ host side
struct NttTables {
  uint32_t *rt = nullptr;  // 29 roots
  uint32_t *tw = nullptr;  // w_(2^Ltab)^k, k < 2^(Ltab-1)
  uint32_t Ltab = 0;
  uint32_t *scal = nullptr;  // per L: 2 scalers
};

static NttTables g_tabs[16];  // per device
static std::mutex g_tabs_mu[16];

// The tables are shared by every context, stream and host thread of a
// device.  ensure_tables hands the caller a snapshot taken under the
// device's lock; a table is never freed while the process runs (a larger
// transform builds a new twiddle table and the old one is retired, not
// released: a launch may still hold it), so a snapshot stays valid.  The
// retired tables are at most as large as the live one together (sizes
// double).  A build finishes before the lock is released.
static int ensure_tables(csk_ctx *ctx, uint32_t L, hipStream_t st, NttTables &snap) {
  const int dev = ctx->device & 15;
  std::lock_guard<std::mutex> lock(g_tabs_mu[dev]);
  NttTables &T = g_tabs[dev];
  if (!T.rt) {
    uint32_t *rt = nullptr, *scal = nullptr;
    CSK_HIP(hipMalloc(&rt, 29 * 32));
    CSK_HIP(hipMalloc(&scal, (NTT_MAX_LG + 2) * 2 * 32));
    k_roots<<<1, 64, 0, st>>>(rt);
    for (uint32_t l = 0; l <= NTT_MAX_LG + 1; l++) k_scalers<<<1, 64, 0, st>>>(l, scal + l * 16);
    CSK_CHECK_LAUNCH();
    CSK_HIP(hipStreamSynchronize(st));
    T.rt = rt;
    T.scal = scal;
  }
  uint32_t want = L < 11 ? 11 : L;
  if (T.Ltab < want) {
    uint32_t *tw = nullptr;
    CSK_HIP(hipMalloc(&tw, ((size_t)1 << (want - 1)) * 32));
    k_twiddles<<<(uint32_t)((((size_t)1 << (want - 1)) + 255) / 256), 256, 0, st>>>(T.rt, want, tw);
    CSK_CHECK_LAUNCH();
    CSK_HIP(hipStreamSynchronize(st));
    T.tw = tw;  // the previous table is retired, not freed
    T.Ltab = want;
  }
  snap = T;
  return CSK_OK;
}

// forward DIF passes src -> dst (in place when equal) for L >= 11; the result
// is in bit-reversed order
static int dif_passes(const uint32_t *src, uint32_t *dst, uint32_t L, const NttTables &T, hipStream_t st) {
  uint32_t npass = (L + 10) / 11;
  // the 1024-element tile when it needs no extra pass (L = 12..20, 23..30)
  const bool small = (L + 9) / 10 == npass;
  uint32_t s_hi = L - 1;
  const uint32_t *in = src;
  for (uint32_t p = 0; p < npass; p++) {
    uint32_t left = s_hi + 1;
    uint32_t R = (left + (npass - p) - 1) / (npass - p);  // spread stages evenly
    uint64_t N = 1ull << L;
    if (small)
      k_ntt_dif_pass<1024><<<(uint32_t)(N / 1024), NTT_THREADS, 0, st>>>(in, dst, L, s_hi, R, T.tw, T.Ltab);
    else
      k_ntt_dif_pass<NTT_TILE><<<(uint32_t)(N / NTT_TILE), NTT_THREADS, 0, st>>>(in, dst, L, s_hi, R, T.tw, T.Ltab);
    CSK_CHECK_LAUNCH();
    in = dst;
    s_hi -= R;
  }
  return CSK_OK;
}

// Transform of N = 2^L elements: src -> dst via work buffer wk (max(N, 2048)
// elements).  inverse: 0 forward, 1 inverse with output fmt (1 Montgomery, 0
// normal).  coset multiplies the result by w_(2N)^i (H-step coset shift);
// mul_by multiplies it pointwise; the result goes to dst[i*ostride + ooff].
// Sizes below one LDS tile run as one pass over a zero-padded tile: the
// transform occupies the first N slots (G = 2048 / N contiguous
// sub-problems), the other sub-problems transform zeros and are not read.
static int ntt_core(csk_ctx *ctx, const uint32_t *src, uint32_t *dst, uint32_t *wk, uint32_t L, int inverse,
                    int fmt, int coset, const uint32_t *mul_by, uint32_t ostride, uint32_t ooff,
                    hipStream_t st) {
  NttTables T;
  int rc = ensure_tables(ctx, L + (coset ? 1 : 0), st, T);
  if (rc) return rc;
  const uint64_t N = 1ull << L;
  if (L < 11) {
    k_fill0<<<NTT_TILE * 8 / 256, 256, 0, st>>>(wk, (uint64_t)NTT_TILE * 8);
    CSK_HIP(hipMemcpyAsync(wk, src, N * 32, hipMemcpyDeviceToDevice, st));
    if (L > 0) k_ntt_dif_pass<NTT_TILE><<<1, NTT_THREADS, 0, st>>>(wk, wk, L, L - 1, L, T.tw, T.Ltab);
    CSK_CHECK_LAUNCH();
  } else {
    rc = dif_passes(src, wk, L, T, st);
    if (rc) return rc;
  }
  const uint32_t *scale = inverse ? T.scal + L * 16 + (fmt ? 0 : 8) : nullptr;
  uint32_t cshift = coset ? (T.Ltab - (L + 1)) : 0;
  k_ntt_finalize<<<(uint32_t)((N + 255) / 256), 256, 0, st>>>(wk, dst, L, inverse ? 1 : 0, scale,
                                                              coset ? T.tw : nullptr, cshift, mul_by, ostride,
                                                              ooff);
  CSK_CHECK_LAUNCH();
  return CSK_OK;
}

}  // namespace csk

using namespace csk;

extern "C" {

int csk_ntt(csk_ctx *ctx, const uint32_t *d_in, uint32_t *d_out, uint32_t log2n, int inverse, int out_fmt,
            void *stream) {
  if (!ctx || !d_in || !d_out || log2n > NTT_MAX_LG) {
    set_error("ntt: bad argument (log2n must be <= 27)");
    return CSK_ERR_ARG;
  }
  CSK_HIP(hipSetDevice(ctx->device));
  uint64_t N = 1ull << log2n;
  int rc = ctx->ws.ensure((N < NTT_TILE ? NTT_TILE : N) * 32);
  if (rc) return rc;
  return ntt_core(ctx, d_in, d_out, (uint32_t *)ctx->ws.ptr, log2n, inverse, out_fmt, 0, nullptr, 1, 0,
                  (hipStream_t)stream);
}

// H step (ntt_interpolandmul): d_out (2m elements, normal form) = coefficients
// of A(x)B(x), from m Montgomery evaluations pA, pB on the m-th roots.
int csk_interpolandmul(csk_ctx *ctx, const uint32_t *d_pA, const uint32_t *d_pB, uint32_t log2m, uint32_t *d_out,
                       void *stream) {
  if (!ctx || !d_pA || !d_pB || !d_out || log2m < 1 || log2m + 1 > NTT_MAX_LG) {
    set_error("interpolandmul: bad argument (2 <= m <= 2^26)");
    return CSK_ERR_ARG;
  }
  CSK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  uint64_t m = 1ull << log2m;
  // workspace: A', B' (m each) + transform scratch (max(2m, one tile))
  uint64_t wkn = 2 * m < (uint64_t)NTT_TILE ? (uint64_t)NTT_TILE : 2 * m;
  int rc = ctx->ws.ensure((2 * m + wkn) * 32);
  if (rc) return rc;
  uint32_t *a = (uint32_t *)ctx->ws.ptr, *b = a + 8 * m, *wk = b + 8 * m;
  uint32_t g = (uint32_t)((m + 255) / 256);
  // even slots: A(w_2m^(2i)) B(w_2m^(2i)) = pA[i] pB[i]
  k_pointwise<<<g, 256, 0, st>>>(d_pA, d_pB, d_out, m, 2, 0);
  CSK_CHECK_LAUNCH();
  // odd slots: interpolate, shift to the coset w_2m <w_m>, evaluate
  rc = ntt_core(ctx, d_pA, a, wk, log2m, 1, 1, 1, nullptr, 1, 0, st);
  if (rc) return rc;
  rc = ntt_core(ctx, a, a, wk, log2m, 0, 1, 0, nullptr, 1, 0, st);
  if (rc) return rc;
  rc = ntt_core(ctx, d_pB, b, wk, log2m, 1, 1, 1, nullptr, 1, 0, st);
  if (rc) return rc;
  rc = ntt_core(ctx, b, d_out, wk, log2m, 0, 1, 0, a, 2, 1, st);  // fused A'[i]*B'[i] -> out[2i+1]
  if (rc) return rc;
  // 2m-point inverse transform, normal-form output
  return ntt_core(ctx, d_out, d_out, wk, log2m + 1, 1, 0, 0, nullptr, 1, 0, st);
}

int csk_mpoly_eval_csr(csk_ctx *ctx, const uint64_t *d_rowptr, const uint32_t *d_col, const uint32_t *d_val,
                       const uint32_t *d_wit, uint32_t nrows, uint32_t *d_out, void *stream) {
  if (!ctx || (nrows && (!d_rowptr || !d_out))) {
    set_error("mpoly_eval_csr: bad argument");
    return CSK_ERR_ARG;
  }
  if (nrows == 0) return CSK_OK;
  CSK_HIP(hipSetDevice(ctx->device));
  k_spmv<<<(nrows + 255) / 256, 256, 0, (hipStream_t)stream>>>(d_rowptr, d_col, d_val, d_wit, nrows, d_out);
  CSK_CHECK_LAUNCH();
  return CSK_OK;
}

}  // extern "C"
