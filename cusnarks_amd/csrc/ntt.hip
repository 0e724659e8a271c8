// Fr number-theoretic transforms, the QAP H step and QAP evaluation on gfx950.
//
// NTT:  X[k] = sum_i a[i] w^(ik), w = 5^((r-1)/N) (cusnarks proots_field,
//       src/cuda/constants.cpp:323-352), natural order in and out, Montgomery
//       data -- the value of ntt_h (src/cuda/ntt.cpp:453-463).
// INTT: N^-1 sum_k X[k] w^(-ik), output Montgomery (fmt 1) or normal (fmt 0),
//       the value of intt_h (ntt.cpp:583-603, IScalerMont/IScalerExt).
//
// Schedule (N >= 2^11): radix-2 decimation-in-frequency stages L-1 .. 0 in P
// passes of R <= 9 stages (k_ntt_r4).  A pass loads 2^R-point sub-transforms
// of G = TILE / 2^R sub-problems into LDS (limb-major, so consecutive lanes
// touch consecutive banks), runs its stages two at a time as radix-4
// butterflies with twiddles from a stage-major table T[2^s + j] = w_(2^(s+1))^j
// (consecutive columns read consecutive entries), and writes back in place.
// Strided passes take G >= 4 adjacent columns, so every row of the tile is a
// full 128-byte line.  The last pass takes G sub-blocks whose bit-reversed
// block numbers are consecutive, so its outputs leave in natural order, G
// adjacent elements per row: the bit reversal costs no gather.  That pass also
// applies the index reversal of the inverse (INTT(X)[i] = N^-1 NTT(X)[-i]), the
// N^-1 scaling, the coset factor and an optional pointwise product.  Smaller
// transforms run as one LDS pass over a zero-padded tile and a finalize gather.
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

// Stage-major table: T[2^s + j] = w_(2^(s+1))^j for j < 2^s, s < Ltab (T[0]
// unused).  w_(2^(s+1))^(2^b) = w_(2^(s+1-b)) = rt[s + 1 - b].  The row of
// stage Ltab - 1, T + 2^(Ltab-1), is the table w_(2^Ltab)^k, k < 2^(Ltab-1),
// that the single-pass kernel and the finalize gather index with a stride.
__global__ void __launch_bounds__(256) k_twiddles(const uint32_t *__restrict__ rt, uint32_t Ltab,
                                                  uint32_t *__restrict__ T) {
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (1ull << Ltab)) return;
  Fr acc = Fr::one();
  if (idx >= 2) {
    const uint32_t st = 63 - __builtin_clzll(idx);
    const uint64_t j = idx - (1ull << st);
    for (uint32_t b = 0; b < st; b++) {
      if ((j >> b) & 1u) {
        Fr f;
        ld8(f.v, rt + 8 * (st + 1 - b));
        acc = Fr::mul(acc, f);
      }
    }
  }
  st8(T + 8 * idx, acc.v);
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

// ------------------------------------------------------------------ radix-4 passes
// What the last pass applies to each output X[k] before storing it: i = k, or
// -k mod N for the inverse; x = X[k] (* scale) (* w_(2N)^i) (* mul_by[i]);
// dst[i * ostride + ooff] = x.
struct NttEpi {
  const uint32_t *scale;   // one element or null
  const uint32_t *coset;   // w_(2N)^i at coset[8 i] (the stage-L row of T) or null
  const uint32_t *mul_by;  // N elements or null
  uint32_t inverse, ostride, ooff;
};

// LDS rows of one limb: position p at word p + p / 64 of its row (one pad
// word per 64), so the strided position sets of the low stages and of the
// bit-reversed epilogue (stride 16 or 2^R words) still fall on distinct banks
__device__ __forceinline__ uint32_t lds_ofs(uint32_t p) { return p + (p >> 6); }
__device__ __forceinline__ Fr lds_ld(const uint32_t *sh, uint32_t rs, uint32_t p) {
  Fr a;
  const uint32_t o = lds_ofs(p);
#pragma unroll
  for (int k = 0; k < 8; k++) a.v[k] = sh[k * rs + o];
  return a;
}
__device__ __forceinline__ void lds_st(uint32_t *sh, uint32_t rs, uint32_t p, const Fr &a) {
  const uint32_t o = lds_ofs(p);
#pragma unroll
  for (int k = 0; k < 8; k++) sh[k * rs + o] = a.v[k];
}

// Batched transforms: one launch runs a pass of up to NTT_BATCH same-size
// transforms, member blockIdx.y.  Each member brings its input, output, work
// buffer (the passes between the first and the last run in place there) and
// optional pointwise factor.  The table travels by value in the kernel
// arguments, so the host needs no device copy of it.
struct NttJob {
  const uint32_t *src;
  uint32_t *dst;
  uint32_t *wk;
  const uint32_t *mul_by;
};
static constexpr uint32_t NTT_BATCH = 32;
struct NttJobs {
  NttJob j[NTT_BATCH];
};
enum NttRole : uint32_t { ROLE_FIRST = 0, ROLE_MID = 1, ROLE_LAST = 2 };

// One pass of stages s_hi .. s_lo = s_hi + 1 - R over N = 2^L elements.
// Tile position (g, t), g < G = TILE / 2^R, t < 2^R:
//   strided pass (s_lo > 0): element (high << (s_hi+1)) | (t << s_lo) | (low0 + g),
//     G adjacent columns low0 .. low0+G-1 (G >= 4: 128-byte rows);
//   last pass (s_lo = 0): element (hi_g << R) | t with hi_g = bitrev_(L-R)(b0 + g),
//     b0 = blockIdx * G; its output X[bitrev_R(t) << (L-R) | (b0 + g)] is written
//     in natural order through the epilogue.
// LDS holds limb k of position p at sh[k * RS + lds_ofs(p)], p = t * G + g
// (strided) or g * 2^R + t (last), so the lanes of a step touch distinct banks.
// Register budget: the 1024-element tile runs three waves per SIMD (142
// VGPRs, no spills; measured faster than four with spills), the 2048-element
// tile two (its 64 KB of LDS allow no more).
// Values between the first load and the last store stay on the lazy range
// [0, 2r) (mul_lazy / add_lazy / sub_lazy: no final subtraction per product);
// the last pass reduces each output once before its epilogue.
// With BATCH the member's buffers come from jobs (src/dst ignored): the
// first pass reads src and writes wk, the middle ones stay in wk, the last
// reads wk and writes dst with the member's mul_by.
template <int TILE, bool LAST, bool BATCH>
__global__ void __launch_bounds__(NTT_THREADS) __attribute__((amdgpu_waves_per_eu(TILE == 2048 ? 2 : 3)))
k_ntt_r4(const uint32_t *__restrict__ src_in, uint32_t *__restrict__ dst_in, uint32_t L, uint32_t s_hi, uint32_t R,
         const uint32_t *__restrict__ T, NttEpi ep,
         typename std::conditional<BATCH, NttJobs, uint32_t>::type jobs, uint32_t role) {
  const uint32_t *src = src_in;
  uint32_t *dst = dst_in;
  if constexpr (BATCH) {
    const NttJob &jb = jobs.j[blockIdx.y];
    src = role == ROLE_FIRST ? jb.src : jb.wk;
    dst = role == ROLE_LAST ? jb.dst : jb.wk;
    ep.mul_by = jb.mul_by;
  }
  constexpr uint32_t RS = TILE + TILE / 64;  // words per limb row
  __shared__ uint32_t sh[8 * RS];
  constexpr uint32_t LGT = TILE == 2048 ? 11 : 10;
  const uint32_t s_lo = s_hi + 1 - R;
  const uint32_t lgG = LGT - R, G = 1u << lgG;
  const uint32_t tmask = (1u << R) - 1;
  uint64_t high = 0;
  uint32_t low0 = 0, b0 = 0;
  if (LAST) {
    b0 = blockIdx.x << lgG;
  } else {
    const uint32_t cg = s_lo - lgG;  // column groups per high value: 2^cg
    low0 = (blockIdx.x & ((1u << cg) - 1)) << lgG;
    high = (uint64_t)(blockIdx.x >> cg) << (s_hi + 1);
  }
  auto pos = [&](uint32_t g, uint32_t t) -> uint32_t { return LAST ? (g << R) | t : (t << lgG) | g; };
  auto elem = [&](uint32_t g, uint32_t t) -> uint64_t {
    if (LAST) return ((uint64_t)(__brev(b0 + g) >> (32 - (L - R))) << R) | t;
    return high | ((uint64_t)t << s_lo) | (low0 + g);
  };
  // load: a lane walks t within one sub-block (last pass) or g across one row
#pragma unroll 4
  for (uint32_t i = threadIdx.x; i < TILE; i += NTT_THREADS) {
    const uint32_t g = LAST ? i >> R : i & (G - 1), t = LAST ? i & tmask : i >> lgG;
    const uint4 *sp = reinterpret_cast<const uint4 *>(src + 8 * elem(g, t));
    const uint4 x = sp[0], y = sp[1];
    const uint32_t p = lds_ofs(pos(g, t));
    sh[0 * RS + p] = x.x; sh[1 * RS + p] = x.y; sh[2 * RS + p] = x.z; sh[3 * RS + p] = x.w;
    sh[4 * RS + p] = y.x; sh[5 * RS + p] = y.y; sh[6 * RS + p] = y.z; sh[7 * RS + p] = y.w;
  }
  __syncthreads();
  // stages on t-bits b = R-1 .. 0, two at a time; an odd R ends with one radix-2 stage
  int b = (int)R - 1;
  for (; b >= 1; b -= 2) {
    const uint32_t h = 1u << (b - 1);
    const uint32_t s = s_lo + b;  // the upper stage of the pair
    const uint32_t qb = R - 2;    // q < 2^(R-2)
#pragma unroll 1
    for (uint32_t idx = threadIdx.x; idx < TILE / 4; idx += NTT_THREADS) {
      const uint32_t g = LAST ? idx >> qb : idx & (G - 1), q = LAST ? idx & ((1u << qb) - 1) : idx >> lgG;
      const uint32_t t0 = ((q >> (b - 1)) << (b + 1)) | (q & (h - 1));
      const uint32_t low = LAST ? 0u : low0 + g;
      const uint64_t jm = ((uint64_t)(t0 & (h - 1)) << s_lo) + low;  // j mod 2^(s-1), also j mod 2^s at t0
      Fr w1, w2, w3;
      ld8(w1.v, T + 8 * ((1ull << s) + jm));
      ld8(w2.v, T + 8 * ((1ull << s) + jm + (1ull << (s - 1))));
      ld8(w3.v, T + 8 * ((1ull << (s - 1)) + jm));
      const uint32_t p0 = pos(g, t0), p1 = pos(g, t0 + h), p2 = pos(g, t0 + 2 * h), p3 = pos(g, t0 + 3 * h);
      const Fr a0 = lds_ld(sh, RS, p0), a1 = lds_ld(sh, RS, p1);
      const Fr a2 = lds_ld(sh, RS, p2), a3 = lds_ld(sh, RS, p3);
      // stage s: (t0, t0+2h) with w1, (t0+h, t0+3h) with w2; the products are
      // kept in sequence (a scheduling barrier after each), which holds the
      // kernel to about a quarter of the registers the interleaved form takes
      const Fr u0 = Fr::add_lazy(a0, a2), u1 = Fr::add_lazy(a1, a3);
      const Fr v0 = Fr::mul_lazy(Fr::sub_lazy(a0, a2), w1);
      __builtin_amdgcn_sched_barrier(0);
      const Fr v1 = Fr::mul_lazy(Fr::sub_lazy(a1, a3), w2);
      __builtin_amdgcn_sched_barrier(0);
      // stage s-1: (t0, t0+h) and (t0+2h, t0+3h), both with w3
      lds_st(sh, RS, p0, Fr::add_lazy(u0, u1));
      lds_st(sh, RS, p2, Fr::add_lazy(v0, v1));
      lds_st(sh, RS, p1, Fr::mul_lazy(Fr::sub_lazy(u0, u1), w3));
      __builtin_amdgcn_sched_barrier(0);
      lds_st(sh, RS, p3, Fr::mul_lazy(Fr::sub_lazy(v0, v1), w3));
    }
    __syncthreads();
  }
  if (b == 0) {  // radix-2 stage s_lo on t-bit 0: pairs (t, t+1)
    const uint32_t qb = R - 1;
#pragma unroll 1
    for (uint32_t idx = threadIdx.x; idx < TILE / 2; idx += NTT_THREADS) {
      const uint32_t g = LAST ? idx >> qb : idx & (G - 1), q = LAST ? idx & ((1u << qb) - 1) : idx >> lgG;
      const uint32_t t = q << 1;
      const uint64_t jm = LAST ? 0u : low0 + g;  // t mod 2^0 = 0
      Fr w;
      ld8(w.v, T + 8 * ((1ull << s_lo) + jm));
      const uint32_t pa = pos(g, t), pc = pos(g, t + 1);
      const Fr a = lds_ld(sh, RS, pa), c = lds_ld(sh, RS, pc);
      lds_st(sh, RS, pa, Fr::add_lazy(a, c));
      lds_st(sh, RS, pc, Fr::mul_lazy(Fr::sub_lazy(a, c), w));
    }
    __syncthreads();
  }
  if (!LAST) {
#pragma unroll 4
    for (uint32_t i = threadIdx.x; i < TILE; i += NTT_THREADS) {
      const uint32_t g = i & (G - 1), t = i >> lgG, p = lds_ofs(pos(g, t));
      uint4 *dp = reinterpret_cast<uint4 *>(dst + 8 * elem(g, t));
      dp[0] = make_uint4(sh[0 * RS + p], sh[1 * RS + p], sh[2 * RS + p], sh[3 * RS + p]);
      dp[1] = make_uint4(sh[4 * RS + p], sh[5 * RS + p], sh[6 * RS + p], sh[7 * RS + p]);
    }
  } else {
    // output row u = bitrev_R(t): G adjacent frequencies (u << (L-R)) | (b0 + g)
    const uint64_t N = 1ull << L;
    Fr sc;
    if (ep.scale) ld8(sc.v, ep.scale);
#pragma unroll 1
    for (uint32_t i = threadIdx.x; i < TILE; i += NTT_THREADS) {
      const uint32_t g = i & (G - 1), u = i >> lgG;
      const uint32_t t = __brev(u) >> (32 - R);
      const uint64_t k = ((uint64_t)u << (L - R)) | (b0 + g);
      const uint64_t io = ep.inverse ? (N - k) & (N - 1) : k;
      Fr x = Fr::reduce_once(lds_ld(sh, RS, pos(g, t))), w = {}, m = {};
      if (ep.coset) ld8(w.v, ep.coset + 8 * io);
      if (ep.mul_by) ld8(m.v, ep.mul_by + 8 * io);
      if (ep.scale) x = Fr::mul(x, sc);
      if (ep.coset) x = Fr::mul(x, w);
      if (ep.mul_by) x = Fr::mul(x, m);
      st8(dst + 8 * (io * ep.ostride + ep.ooff), x.v);
    }
  }
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

// batched pointwise products: member y, out[2i] = src[i] * mul_by[i]
__global__ void __launch_bounds__(256) k_pointwise_b(NttJobs jobs, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const NttJob &jb = jobs.j[blockIdx.y];
  Fr x, y;
  ld8(x.v, jb.src + 8 * i);
  ld8(y.v, jb.mul_by + 8 * i);
  st8(jb.dst + 8 * (2 * i), Fr::mul(x, y).v);
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
  uint32_t *tw = nullptr;  // stage-major twiddles T[2^s + j] = w_(2^(s+1))^j, 2^Ltab entries
  uint32_t Ltab = 0;
  uint32_t *scal = nullptr;  // per L: 2 scalers
  // w_(2^Ltab)^k for k < 2^(Ltab-1): the top row of the stage-major table
  const uint32_t *row() const { return tw + 8 * ((size_t)1 << (Ltab - 1)); }
};

static NttTables g_tabs[16];  // per device
static std::mutex g_tabs_mu[16];

// The tables are shared by every context, stream and host thread of a
// device.  ensure_tables hands the caller a snapshot taken under the
// device's lock; a table is never freed while the process runs (a larger
// transform builds a new twiddle table and the old one is retired, not
// released: a launch may still hold it), so a snapshot stays valid.  The
// retired tables are at most as large as the live one together (sizes
// double).  A build finishes before the lock is released.  2^27 entries
// (4 GiB) serve transforms up to 2^26 points with the coset row.
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
    CSK_HIP(hipMalloc(&tw, ((size_t)1 << want) * 32));
    k_twiddles<<<(uint32_t)((((size_t)1 << want) + 255) / 256), 256, 0, st>>>(T.rt, want, tw);
    CSK_CHECK_LAUNCH();
    CSK_HIP(hipStreamSynchronize(st));
    T.tw = tw;  // the previous table is retired, not freed
    T.Ltab = want;
  }
  snap = T;
  return CSK_OK;
}

// Radix-4 passes for L >= 11: src -> wk (first pass), in place on wk, then wk
// -> dst in natural order through the epilogue (last pass).  Passes take at
// most RMAX stages so a strided pass keeps G >= 4 columns per tile; the stage
// counts are spread evenly, the larger ones first.
// 1024-element tiles (32 KB of LDS): H step at m = 2^25 33.7 ms against 36.4
// ms with 2048-element tiles (two workgroups per CU instead of five)
static constexpr int R4_TILE = 1024;

template <int TILE>
static int r4_passes(const uint32_t *src, uint32_t *wk, uint32_t *dst, uint32_t L, const NttTables &T,
                     const NttEpi &ep, hipStream_t st) {
  constexpr uint32_t R4_LGT = TILE == 2048 ? 11 : 10, R4_RMAX = TILE == 2048 ? 9 : 8;
  const uint32_t P = (L + R4_RMAX - 1) / R4_RMAX;
  const uint32_t grid = (uint32_t)((1ull << L) >> R4_LGT);
  uint32_t s_hi = L - 1;
  const uint32_t *in = src;
  for (uint32_t p = 0; p < P; p++) {
    const uint32_t left = s_hi + 1;
    const uint32_t R = (left + (P - p) - 1) / (P - p);
    const uint32_t s_lo = left - R, lgG = R4_LGT - R;
    if (R > R4_RMAX || R < 2 || (p + 1 < P ? s_lo < lgG : L - R < lgG)) {
      set_error("ntt: no pass plan for this size");
      return CSK_ERR_ARG;
    }
    if (p + 1 < P) {
      k_ntt_r4<TILE, false, false><<<grid, NTT_THREADS, 0, st>>>(in, wk, L, s_hi, R, T.tw, ep, 0u, 0u);
      in = wk;
      s_hi -= R;
    } else {
      k_ntt_r4<TILE, true, false><<<grid, NTT_THREADS, 0, st>>>(in, dst, L, s_hi, R, T.tw, ep, 0u, 0u);
    }
    CSK_CHECK_LAUNCH();
  }
  return CSK_OK;
}

// The same passes for `count` same-size members (count <= NTT_BATCH).
template <int TILE>
static int r4_passes_batch(const NttJobs &jobs, uint32_t count, uint32_t L, const NttTables &T, const NttEpi &ep,
                           hipStream_t st) {
  constexpr uint32_t R4_LGT = TILE == 2048 ? 11 : 10, R4_RMAX = TILE == 2048 ? 9 : 8;
  const uint32_t P = (L + R4_RMAX - 1) / R4_RMAX;
  const dim3 grid((uint32_t)((1ull << L) >> R4_LGT), count);
  uint32_t s_hi = L - 1;
  for (uint32_t p = 0; p < P; p++) {
    const uint32_t left = s_hi + 1;
    const uint32_t R = (left + (P - p) - 1) / (P - p);
    const uint32_t s_lo = left - R, lgG = R4_LGT - R;
    if (R > R4_RMAX || R < 2 || P < 2 || (p + 1 < P ? s_lo < lgG : L - R < lgG)) {
      set_error("ntt: no pass plan for this size");
      return CSK_ERR_ARG;
    }
    if (p + 1 < P) {
      k_ntt_r4<TILE, false, true><<<grid, NTT_THREADS, 0, st>>>(nullptr, nullptr, L, s_hi, R, T.tw, ep, jobs,
                                                                p == 0 ? ROLE_FIRST : ROLE_MID);
      s_hi -= R;
    } else {
      k_ntt_r4<TILE, true, true><<<grid, NTT_THREADS, 0, st>>>(nullptr, nullptr, L, s_hi, R, T.tw, ep, jobs,
                                                               ROLE_LAST);
    }
    CSK_CHECK_LAUNCH();
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
  const uint32_t *scale = inverse ? T.scal + L * 16 + (fmt ? 0 : 8) : nullptr;
  if (L >= 11) {
    NttEpi ep{scale, coset ? T.tw + 8 * N : nullptr, mul_by, (uint32_t)(inverse ? 1 : 0), ostride, ooff};
    return r4_passes<R4_TILE>(src, wk, dst, L, T, ep, st);
  }
  k_fill0<<<NTT_TILE * 8 / 256, 256, 0, st>>>(wk, (uint64_t)NTT_TILE * 8);
  CSK_HIP(hipMemcpyAsync(wk, src, N * 32, hipMemcpyDeviceToDevice, st));
  if (L > 0) k_ntt_dif_pass<NTT_TILE><<<1, NTT_THREADS, 0, st>>>(wk, wk, L, L - 1, L, T.row(), T.Ltab);
  CSK_CHECK_LAUNCH();
  uint32_t cshift = coset ? (T.Ltab - (L + 1)) : 0;
  k_ntt_finalize<<<(uint32_t)((N + 255) / 256), 256, 0, st>>>(wk, dst, L, inverse ? 1 : 0, scale,
                                                              coset ? T.row() : nullptr, cshift, mul_by, ostride,
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

// H steps of `count` same-size problems (member t: d_pA[t], d_pB[t] ->
// d_out[t]) with every transform pass as one launch over up to NTT_BATCH
// members: the results equal csk_interpolandmul's member by member.
int csk_interpolandmul_batch(csk_ctx *ctx, uint32_t count, const uint32_t *const *d_pA,
                             const uint32_t *const *d_pB, uint32_t log2m, uint32_t *const *d_out, void *stream) {
  if (!ctx || (count && (!d_pA || !d_pB || !d_out)) || log2m < 1 || log2m + 1 > NTT_MAX_LG) {
    set_error("interpolandmul_batch: bad argument (2 <= m <= 2^26)");
    return CSK_ERR_ARG;
  }
  for (uint32_t t = 0; t < count; t++)
    if (!d_pA[t] || !d_pB[t] || !d_out[t]) {
      set_error("interpolandmul_batch: null member buffer");
      return CSK_ERR_ARG;
    }
  if (count == 0) return CSK_OK;
  if (log2m < 11) {  // below the radix-4 passes: member by member
    for (uint32_t t = 0; t < count; t++) {
      int rc = csk_interpolandmul(ctx, d_pA[t], d_pB[t], log2m, d_out[t], stream);
      if (rc) return rc;
    }
    return CSK_OK;
  }
  CSK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  const uint64_t m = 1ull << log2m;
  const uint32_t chunk = count < NTT_BATCH ? count : NTT_BATCH;
  // per member: A', B' (m each) and the transform work buffer (2m)
  int rc = ctx->ws.ensure((size_t)chunk * 4 * m * 32);
  if (rc) return rc;
  NttTables T;
  rc = ensure_tables(ctx, log2m + 1, st, T);
  if (rc) return rc;
  uint32_t *base = (uint32_t *)ctx->ws.ptr;
  const uint32_t *scale_m = T.scal + log2m * 16;             // N^-1 R: Montgomery output
  const uint32_t *scale_2m = T.scal + (log2m + 1) * 16 + 8;  // N^-1: normal output
  const uint32_t *coset = T.tw + 8 * m;                      // w_2m^i
  for (uint32_t c0 = 0; c0 < count; c0 += NTT_BATCH) {
    const uint32_t k = count - c0 < NTT_BATCH ? count - c0 : NTT_BATCH;
    NttJobs ev = {}, ia = {}, fa = {}, ib = {}, fb = {}, i2 = {};
    for (uint32_t t = 0; t < k; t++) {
      uint32_t *a = base + (uint64_t)t * 4 * m * 8, *b = a + m * 8, *wk = b + m * 8;
      uint32_t *out = d_out[c0 + t];
      ev.j[t] = {d_pA[c0 + t], out, nullptr, d_pB[c0 + t]};
      ia.j[t] = {d_pA[c0 + t], a, wk, nullptr};
      fa.j[t] = {a, a, wk, nullptr};
      ib.j[t] = {d_pB[c0 + t], b, wk, nullptr};
      fb.j[t] = {b, out, wk, a};
      i2.j[t] = {out, out, wk, nullptr};
    }
    // even slots: A(w_2m^(2i)) B(w_2m^(2i)) = pA[i] pB[i]
    k_pointwise_b<<<dim3((uint32_t)((m + 255) / 256), k), 256, 0, st>>>(ev, m);
    CSK_CHECK_LAUNCH();
    // odd slots: interpolate, shift to the coset w_2m <w_m>, evaluate, multiply
    if ((rc = r4_passes_batch<R4_TILE>(ia, k, log2m, T, NttEpi{scale_m, coset, nullptr, 1, 1, 0}, st))) return rc;
    if ((rc = r4_passes_batch<R4_TILE>(fa, k, log2m, T, NttEpi{nullptr, nullptr, nullptr, 0, 1, 0}, st))) return rc;
    if ((rc = r4_passes_batch<R4_TILE>(ib, k, log2m, T, NttEpi{scale_m, coset, nullptr, 1, 1, 0}, st))) return rc;
    if ((rc = r4_passes_batch<R4_TILE>(fb, k, log2m, T, NttEpi{nullptr, nullptr, nullptr, 0, 2, 1}, st))) return rc;
    // 2m-point inverse transform, normal-form output
    if ((rc = r4_passes_batch<R4_TILE>(i2, k, log2m + 1, T, NttEpi{scale_2m, nullptr, nullptr, 1, 1, 0}, st)))
      return rc;
  }
  return CSK_OK;
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
