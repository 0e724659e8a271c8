// Pippenger kernels and the per-curve launch sequence (see msm.hip for the
// pipeline overview).  Included by msm_g1.hip and msm_g2.hip so that the two
// curves compile as separate translation units.
#pragma once
#include <type_traits>
#include "internal.h"
#include "curve.h"
#include "field29.h"
#include "coop29.h"
#include "msm.h"

namespace csk {


__constant__ uint32_t kR1[8] = {4026531841u, 1138881939u, 2042196113u, 674490440u,
                                2172737629u, 3092268470u, 3778125865u, 811880050u};
__constant__ uint32_t kR2[8] = {3758096386u, 2277763879u, 4084392226u, 1348980880u,
                                50507962u,   1889569645u, 3261284435u, 1623760101u};
__constant__ uint32_t kR4[8] = {3221225476u, 260560463u,  3873817157u, 2697961761u,
                                101015924u,  3779139290u, 2227601574u, 3247520203u};

__device__ __forceinline__ void cond_sub(uint32_t *k, const uint32_t *m) {
  uint32_t d[8], borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    uint64_t t = (uint64_t)k[i] - m[i] - borrow;
    d[i] = (uint32_t)t;
    borrow = (uint32_t)(t >> 63);
  }
#pragma unroll
  for (int i = 0; i < 8; i++) k[i] = borrow ? k[i] : d[i];
}

// k < 2^256 -> k mod r (2^256 < 6r, so three conditional subtractions)
__device__ __forceinline__ void reduce_mod_r(uint32_t *k) {
  cond_sub(k, kR4);
  cond_sub(k, kR2);
  cond_sub(k, kR1);
}

// ------------------------------------------------------------------ members
// largest j < J with segs[j].pbase <= g (the member holding concatenated point
// g; empty members share their pbase with the next one and are skipped)
__device__ __forceinline__ uint32_t member_of_point(const MsmSeg *segs, uint32_t J, uint64_t g) {
  uint32_t lo = 0, hi = J;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (segs[mid].pbase <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// largest j < J with segs[j].tbase <= b (the member of pass-1 workgroup b)
__device__ __forceinline__ uint32_t member_of_tile(const MsmSeg *segs, uint32_t J, uint32_t b) {
  uint32_t lo = 0, hi = J;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (segs[mid].tbase <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

// the point in the 2^261 form (field29.h), packed words, component by component
__device__ __forceinline__ void store_point29(uint32_t *d, const Affine<Fq> &p) {
  uint32_t o[8];
  to29_words(p.x.v, o);
  st8(d, o);
  to29_words(p.y.v, o);
  st8(d + 8, o);
}
__device__ __forceinline__ void store_point29(uint32_t *d, const Affine<Fq2> &p) {
  uint32_t o[8];
  to29_words(p.x.c0.v, o);
  st8(d, o);
  to29_words(p.x.c1.v, o);
  st8(d + 8, o);
  to29_words(p.y.c0.v, o);
  st8(d + 16, o);
  to29_words(p.y.c1.v, o);
  st8(d + 24, o);
}

// ------------------------------------------------------------------ digits
// One workgroup per (member j, tile) of kSortTile points; block b covers the
// tiles in member order (member j's tiles start at b = tbase_j / W).  Each
// thread takes every 256th point of the tile and writes its signed digits:
// member j's digits occupy D[W*pbase_j + w*n_j + i], b | neg << 31 for the
// nonzero signed digit d of window w (b = |d| - 1 < B), or 0xFFFFFFFF when
// d == 0 or P_i is at infinity.  The workgroup also counts its digits per
// (window, hi = b >> s) in LDS and writes the counts as pass 1 of the sort
// expects them: cnt[cbase_j + (w*HB + hi)*NT_j + tile].  It also writes each
// point in the 2^261 form for the bucket loop, at concatenated point number
// pbase_j + i of pts29.
template <class F>
__global__ void __launch_bounds__(256) k_digits(const MsmSeg *__restrict__ segs, uint32_t J, uint32_t c,
                                                uint32_t W, uint32_t s, uint32_t ts, uint32_t HB,
                                                uint32_t *__restrict__ digits, uint32_t *__restrict__ cnt,
                                                uint32_t *__restrict__ pts29) {
  extern __shared__ uint32_t hist[];  // W * HB counters
  const uint32_t j = member_of_tile(segs, J, W * blockIdx.x);
  const uint32_t tile = blockIdx.x - segs[j].tbase / W;
  const uint64_t pbase = segs[j].pbase, n = segs[j].n;
  const uint32_t NT = segs[j].NT;
  for (uint32_t i = threadIdx.x; i < W * HB; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  constexpr int PW = 2 * FieldIO<F>::NW;
  const uint32_t B = 1u << (c - 1);
  const uint32_t mask = (1u << c) - 1;
  uint32_t *dj = digits + W * pbase;
  for (uint32_t t = threadIdx.x; t < kSortTile; t += blockDim.x) {
    const uint64_t i = (uint64_t)tile * kSortTile + t;
    if (i >= n) break;
    uint32_t k[9];
    ld8(k, segs[j].scalars + i * 8);
    k[8] = 0;
    reduce_mod_r(k);
    const Affine<F> p = load_affine<F>(segs[j].points + i * PW);
    const bool zero = affine_is_inf(p);
    store_point29(pts29 + (pbase + i) * PW, p);
    uint32_t carry = 0;
    for (uint32_t w = 0; w < W; w++) {
      uint32_t off = w * c;
      uint32_t li = off >> 5, sh = off & 31;
      uint64_t two = (uint64_t)k[li] | ((uint64_t)(li < 7 ? k[li + 1] : 0u) << 32);
      uint32_t raw = (uint32_t)(two >> sh) & mask;
      raw += carry;
      uint32_t mag, neg;
      if (raw > B) {
        mag = (1u << c) - raw;
        neg = 1;
        carry = 1;
      } else {
        mag = raw;
        neg = 0;
        carry = 0;
      }
      const bool live = mag != 0 && !zero;
      dj[(uint64_t)w * n + i] = live ? ((mag - 1) | (neg << 31)) : 0xFFFFFFFFu;
      if (live) atomicAdd(&hist[w * HB + ((mag - 1) >> (w == W - 1 ? ts : s))], 1u);
    }
  }
  __syncthreads();
  const uint32_t cb = segs[j].cbase + tile;
  for (uint32_t i = threadIdx.x; i < W * HB; i += blockDim.x) cnt[cb + (uint64_t)i * NT] = hist[i];
}

namespace {  // curve-independent kernels: one private copy per translation unit

// ------------------------------------------------------------------ scan
static constexpr int SCAN_THREADS = 256;
static constexpr int SCAN_ITEMS = 8;
static constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const uint32_t *__restrict__ in, uint32_t K,
                                                              uint32_t *__restrict__ bsum) {
  __shared__ uint32_t sh[SCAN_THREADS];
  uint32_t base = blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
  for (int j = 0; j < SCAN_ITEMS; j++) {
    uint32_t idx = base + j * SCAN_THREADS + threadIdx.x;
    if (idx < K) s += in[idx];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = SCAN_THREADS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

// exclusive scan of the block sums in place (any nb)
__global__ void __launch_bounds__(1024) k_scan_blocks(uint32_t *__restrict__ bsum, uint32_t nb) {
  __shared__ uint32_t sh[1024];
  const uint32_t per = (nb + 1023) / 1024;
  uint32_t lo = threadIdx.x * per, s = 0;
  for (uint32_t j = 0; j < per; j++)
    if (lo + j < nb) s += bsum[lo + j];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 1; st < 1024; st <<= 1) {
    uint32_t v = threadIdx.x >= (uint32_t)st ? sh[threadIdx.x - st] : 0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = sh[threadIdx.x] - s;  // exclusive prefix of this thread's run
  for (uint32_t j = 0; j < per; j++)
    if (lo + j < nb) {
      uint32_t v = bsum[lo + j];
      bsum[lo + j] = run;
      run += v;
    }
}

// out[0..K] = exclusive prefix sums of in[0..K) (out[K] = total); in == out allowed
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_final(const uint32_t *in, uint32_t K,
                                                             const uint32_t *__restrict__ bsum, uint32_t *out) {
  __shared__ uint32_t sh[SCAN_THREADS];
  uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;  // contiguous items per thread
  uint32_t v[SCAN_ITEMS], s = 0;
  for (int j = 0; j < SCAN_ITEMS; j++) {
    uint32_t idx = base + j;
    v[j] = idx < K ? in[idx] : 0;
    s += v[j];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 1; st < SCAN_THREADS; st <<= 1) {
    uint32_t t = threadIdx.x >= (uint32_t)st ? sh[threadIdx.x - st] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = bsum[blockIdx.x] + sh[threadIdx.x] - s;
  for (int j = 0; j < SCAN_ITEMS; j++) {
    uint32_t idx = base + j;
    if (idx < K) out[idx] = run;
    run += v[j];
    if (idx + 1 == K) out[K] = run;
  }
}

// exclusive scan of arr[0..len) in LDS by one workgroup of SORT_THREADS;
// tmp holds SORT_THREADS / 64 words (one per wave); returns the total (all threads)
static constexpr int SORT_THREADS = 256;
static constexpr int SORT_TILE = 4096;  // digits of one window per pass-1 workgroup
static constexpr int SORT_IPT = SORT_TILE / SORT_THREADS;

__device__ uint32_t block_exclusive_scan(uint32_t *arr, uint32_t len, uint32_t *tmp) {
  const uint32_t per = (len + SORT_THREADS - 1) / SORT_THREADS;
  const uint32_t lo = threadIdx.x * per;
  uint32_t s = 0;
  for (uint32_t j = 0; j < per; j++)
    if (lo + j < len) s += arr[lo + j];
  // inclusive scan across the 64 lanes of the wave, then the wave totals
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t x = s;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) tmp[wv] = x;
  __syncthreads();
  uint32_t run = x - s, total = 0;
  for (uint32_t k = 0; k < SORT_THREADS / 64; k++) {
    const uint32_t t = tmp[k];
    run += k < wv ? t : 0;
    total += t;
  }
  for (uint32_t j = 0; j < per; j++)
    if (lo + j < len) {
      uint32_t v = arr[lo + j];
      arr[lo + j] = run;
      run += v;
    }
  __syncthreads();
  return total;
}

// ------------------------------------------------------------------ sort pass 1
// Bucket b = hi * 2^s + lo.  Pass 1 groups the entries of every window by hi
// (HB <= 256 bins).  Workgroup x of the group maps to (member j, window w,
// tile) with x = tbase_j + w * NT_j + tile.  k_digits has counted the tile's
// kSortTile digits into cnt[cbase_j + (w*HB + hi)*NT_j + tile]; after an
// exclusive scan of cnt (member-, window-, bin-major), each (bin, tile) owns a
// contiguous output range, and the runs appear in bucket-key order
// j*K + w*B + b.
struct TileRef {
  uint32_t j, w, tile, NT;
  uint64_t n;
  const uint32_t *Dw;
  uint32_t cb;  // counter index of (w, bin 0, tile)
};

__device__ __forceinline__ TileRef tile_ref(const MsmSeg *segs, uint32_t J, uint32_t HB, uint32_t W,
                                            const uint32_t *D) {
  TileRef t;
  t.j = member_of_tile(segs, J, blockIdx.x);
  t.NT = segs[t.j].NT;
  const uint32_t local = blockIdx.x - segs[t.j].tbase;
  t.w = local / t.NT;
  t.tile = local - t.w * t.NT;
  t.n = segs[t.j].n;
  t.Dw = D + W * segs[t.j].pbase + (uint64_t)t.w * t.n;
  t.cb = segs[t.j].cbase + t.w * HB * t.NT + t.tile;
  return t;
}

// Pass 1 placement: the workgroup counting-sorts its entries by hi in LDS and
// writes each bin's run contiguously to the pass-1 arrays: tix (point index |
// sign) and tlo (lo, the low s < 16 bits of the bucket).  The top window of
// every scalar (global window index gw = j * W + w with gw mod Wwin = Wwin - 1:
// w = W - 1 for multiexp members, every Wwin-th member for the one-window
// members of table runs) splits its buckets at ts bits instead; with ts = 0
// its bins are single buckets, the pass-1 ranges are the final ones, and the
// entries go straight to `sorted`.
// HOIST issues the tile's sixteen digit loads before the counts (unconditional,
// the index clamped into the member's row; the tile's first index is < n);
// otherwise each entry is one guarded load, wait and count.  Measured: HOIST
// is faster for groups below ~2^27 entries (latency-bound), slower above
// (2^24 points x 15 windows: 1.71 vs 1.37 ms), where the loads of the many
// resident workgroups already keep HBM busy.
template <bool HOIST>
__global__ void __launch_bounds__(SORT_THREADS) k_sort_scatter1(const uint32_t *__restrict__ D,
                                                                const MsmSeg *__restrict__ segs, uint32_t J,
                                                                uint32_t s, uint32_t ts, uint32_t HB, uint32_t W,
                                                                uint32_t Wwin, const uint32_t *__restrict__ goff,
                                                                uint32_t *__restrict__ tix,
                                                                uint16_t *__restrict__ tlo,
                                                                uint32_t *__restrict__ sorted) {
  __shared__ uint32_t h[256];
  __shared__ uint32_t base[256];
  __shared__ uint32_t scan_tmp[SORT_THREADS];
  __shared__ uint32_t stage[SORT_TILE];
  __shared__ uint16_t slo[SORT_TILE];
  __shared__ uint16_t sbin[SORT_TILE];
  const TileRef t = tile_ref(segs, J, HB, W, D);
  if ((t.j * W + t.w) % Wwin == Wwin - 1) s = ts;
  const uint32_t lomask = (1u << s) - 1;
  for (uint32_t i = threadIdx.x; i < HB; i += SORT_THREADS) h[i] = 0;
  __syncthreads();
  uint32_t dv[SORT_IPT], rk[SORT_IPT];
  if constexpr (HOIST) {
    for (int j = 0; j < SORT_IPT; j++) {
      const uint64_t idx = (uint64_t)t.tile * SORT_TILE + j * SORT_THREADS + threadIdx.x;
      dv[j] = t.Dw[idx < t.n ? idx : t.n - 1];
    }
    for (int j = 0; j < SORT_IPT; j++) {
      const uint64_t idx = (uint64_t)t.tile * SORT_TILE + j * SORT_THREADS + threadIdx.x;
      if (idx >= t.n) dv[j] = 0xFFFFFFFFu;
      const uint32_t d = dv[j];
      rk[j] = d != 0xFFFFFFFFu ? atomicAdd(&h[(d & 0x7FFFFFFFu) >> s], 1u) : 0u;
    }
  } else {
    for (int j = 0; j < SORT_IPT; j++) {
      const uint64_t idx = (uint64_t)t.tile * SORT_TILE + j * SORT_THREADS + threadIdx.x;
      const uint32_t d = idx < t.n ? t.Dw[idx] : 0xFFFFFFFFu;
      dv[j] = d;
      rk[j] = d != 0xFFFFFFFFu ? atomicAdd(&h[(d & 0x7FFFFFFFu) >> s], 1u) : 0u;
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < HB; i += SORT_THREADS) base[i] = h[i];
  __syncthreads();
  const uint32_t total = block_exclusive_scan(base, HB, scan_tmp);
  // h (the counts are in base now) takes each bin's output start minus its
  // tile-local start; offsets are < 2^31, the u32 difference wraps back on +p
  for (uint32_t i = threadIdx.x; i < HB; i += SORT_THREADS) h[i] = goff[t.cb + i * t.NT] - base[i];
  for (int j = 0; j < SORT_IPT; j++) {
    uint32_t d = dv[j];
    if (d == 0xFFFFFFFFu) continue;
    uint32_t b = d & 0x7FFFFFFFu, bin = b >> s;
    uint32_t pos = base[bin] + rk[j];
    uint32_t idx = t.tile * SORT_TILE + j * SORT_THREADS + threadIdx.x;
    stage[pos] = idx | (d & 0x80000000u);
    slo[pos] = (uint16_t)(b & lomask);
    sbin[pos] = (uint16_t)bin;
  }
  __syncthreads();
  if (s == 0) {
    for (uint32_t p = threadIdx.x; p < total; p += SORT_THREADS) sorted[h[sbin[p]] + p] = stage[p];
    return;
  }
  for (uint32_t p = threadIdx.x; p < total; p += SORT_THREADS) {
    const uint32_t g = h[sbin[p]] + p;
    tix[g] = stage[p];
    tlo[g] = slo[p];
  }
}

// ------------------------------------------------------------------ sort pass 2
// One workgroup per (member, window, hi) segment of the pass-1 arrays, seg = (j*W + w)*HB
// + hi: partition by lo into the 2^s buckets seg*2^s + lo (= j*K + w*B +
// hi*2^s + lo).  The segment is counted once (bucket offsets are written for
// every bucket), then placed TILE entries at a time: each tile is
// counting-sorted by lo in LDS and every bucket's run is written contiguously
// after that bucket's running cursor.  TILE entries per step: 8192 measured
// faster for groups of 2^27 entries and more (2^24 x 15 windows), 4096 below.
// A top window (see pass 1) partitions by its ts low bits into the buckets
// gw*B + hi*2^ts + lo; its last segment also points the buckets it cannot
// reach (from HB*2^ts on) at its end, and with ts = 0 pass 1 has placed its
// entries already, so only the offsets are written.
template <int TILE>
__global__ void __launch_bounds__(SORT_THREADS) k_sort_pass2(const uint32_t *__restrict__ tix,
                                                             const uint16_t *__restrict__ tlo,
                                                             const uint32_t *__restrict__ goff,
                                                             const MsmSeg *__restrict__ segs, uint32_t HB,
                                                             uint32_t s, uint32_t ts, uint32_t W, uint32_t Wwin,
                                                             uint32_t Ktot, uint32_t *__restrict__ offsets,
                                                             uint32_t *__restrict__ sorted) {
  extern __shared__ uint32_t dyn[];  // 3 * 2^s words: counts, cursors, tile bases
  __shared__ uint32_t scan_tmp[SORT_THREADS];
  __shared__ uint32_t stage[TILE];
  __shared__ uint16_t sbin[TILE];
  const uint32_t seg = blockIdx.x, per = W * HB, mj = seg / per, rem = seg - mj * per;
  const uint32_t NT = segs[mj].NT, cb = segs[mj].cbase;
  const uint32_t w = rem / HB, hi = rem - w * HB, gw = mj * W + w, B = HB << s;
  if (gw % Wwin == Wwin - 1) s = ts;
  const uint32_t L2 = 1u << s, key0 = gw * B + hi * L2;
  uint32_t *cntv = dyn, *cur = dyn + L2, *tb = dyn + 2 * L2;
  const uint32_t start = goff[cb + rem * NT], end = goff[cb + (rem + 1) * NT];
  if (seg == gridDim.x - 1 && threadIdx.x == 0) offsets[Ktot] = end;
  if (hi == HB - 1)
    for (uint32_t k = key0 + L2 + threadIdx.x; k < (gw + 1) * B; k += SORT_THREADS) offsets[k] = end;
  if (s == 0) {
    if (threadIdx.x == 0) offsets[key0] = start;
    return;
  }
  for (uint32_t i = threadIdx.x; i < L2; i += SORT_THREADS) cntv[i] = 0;
  __syncthreads();
  // eight unconditional (clamped) loads in flight per thread, then their counts
  for (uint32_t e0 = start; e0 < end; e0 += 8 * SORT_THREADS) {
    uint32_t l[8];
    for (int k = 0; k < 8; k++) {
      const uint32_t e = e0 + k * SORT_THREADS + threadIdx.x;
      l[k] = tlo[e < end ? e : end - 1];
    }
    for (int k = 0; k < 8; k++)
      if (e0 + k * SORT_THREADS + threadIdx.x < end) atomicAdd(&cntv[l[k]], 1u);
  }
  __syncthreads();
  block_exclusive_scan(cntv, L2, scan_tmp);
  for (uint32_t i = threadIdx.x; i < L2; i += SORT_THREADS) {
    offsets[key0 + i] = start + cntv[i];
    cur[i] = start + cntv[i];
  }
  __syncthreads();
  for (uint32_t t0 = start; t0 < end; t0 += TILE) {
    const uint32_t tn = min((uint32_t)TILE, end - t0);
    for (uint32_t i = threadIdx.x; i < L2; i += SORT_THREADS) tb[i] = 0;
    __syncthreads();
    constexpr int IPT = TILE / SORT_THREADS;
    uint32_t v[IPT], vl[IPT], rk[IPT];
    for (int j = 0; j < IPT; j++) {  // unconditional loads (clamped; tn >= 1)
      const uint32_t q = j * SORT_THREADS + threadIdx.x, qc = q < tn ? q : tn - 1;
      v[j] = tix[t0 + qc];
      vl[j] = tlo[t0 + qc];
    }
    for (int j = 0; j < IPT; j++) {
      const uint32_t q = j * SORT_THREADS + threadIdx.x;
      rk[j] = q < tn ? atomicAdd(&tb[vl[j]], 1u) : 0u;
    }
    __syncthreads();
    // tile-local counts -> tile-local bases; keep the counts for the cursor update
    for (uint32_t i = threadIdx.x; i < L2; i += SORT_THREADS) cntv[i] = tb[i];
    __syncthreads();
    block_exclusive_scan(tb, L2, scan_tmp);
    for (int j = 0; j < IPT; j++) {
      uint32_t q = j * SORT_THREADS + threadIdx.x;
      if (q < tn) {
        uint32_t lo = vl[j];
        uint32_t pos = tb[lo] + rk[j];
        stage[pos] = v[j];
        sbin[pos] = (uint16_t)lo;
      }
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < tn; p += SORT_THREADS) {
      uint32_t lo = sbin[p];
      sorted[cur[lo] + (p - tb[lo])] = stage[p];
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < L2; i += SORT_THREADS) cur[i] += cntv[i];
    __syncthreads();
  }
}

// largest key in [0, K) with off[key] <= e (requires off[0] <= e < off[K])
__device__ __forceinline__ uint32_t find_key(const uint32_t *__restrict__ off, uint32_t K, uint32_t e) {
  uint32_t lo = 0, hi = K;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The bucket holding entry e, from a key k with off[k] <= e (requires
// e < off[K]): one load when bucket k holds e.  Runs of empty buckets are
// crossed by doubling steps, then bisection: a top window reaches only
// 2^(254 mod c) of its 2^(c-1) buckets (c = 20: 2^14 of 2^19), and where the
// next member's keys follow, stepping over the rest one key at a time held
// the whole launch on one thread (four 20-bit table runs: +27 ms).
__device__ __forceinline__ uint32_t key_from(const uint32_t *__restrict__ off, uint32_t K, uint32_t k, uint32_t e) {
  if (off[k + 1] > e) return k;
  uint32_t lo = k + 1, hi = K, step = 1;  // off[lo] <= e < off[hi]
  for (;;) {
    const uint64_t nx = (uint64_t)lo + step;
    if (nx >= K) break;
    if (off[nx] > e) {
      hi = (uint32_t)nx;
      break;
    }
    lo = (uint32_t)nx;
    step <<= 1;
  }
  while (hi - lo > 1) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (off[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

}  // namespace

// Sequential reader of one thread's run of the sorted (point index | sign)
// array, four entries per 16-byte load: every lane walks its own run, and
// with random point loads streaming through L2 a line read 4 bytes at a time
// is evicted and fetched again before its 32 entries are used.  The array
// is padded by 4 entries so the last aligned load stays inside it.
struct IndexReader {
  const uint32_t *base;
  uint32_t blk;  // e & ~3 of the loaded quad
  uint4 q;
  __device__ __forceinline__ IndexReader(const uint32_t *sorted, uint32_t e) : base(sorted), blk(e & ~3u) {
    q = *reinterpret_cast<const uint4 *>(base + blk);
  }
  __device__ __forceinline__ uint32_t at(uint32_t e) {
    if ((e & ~3u) != blk) {
      blk = e & ~3u;
      q = *reinterpret_cast<const uint4 *>(base + blk);
    }
    const uint32_t k = e & 3u;
    return k == 0 ? q.x : k == 1 ? q.y : k == 2 ? q.z : q.w;
  }
};

// ------------------------------------------------------------ accumulate
// The bucket loops run on 29-bit limbs (field29.h).  k_digits has written the
// group's points in the 2^261 form (packed words, at concatenated point
// number g) at the head of the sort buffer.
//
// Thread t of k_accumulate29 folds sorted entries [t*L, min((t+1)*L, E)) of
// the group.  Every lane performs exactly one mixed addition per loop trip
// (lane-uniform trip count), and a run is flushed when the entry just added
// closes its bucket or the chunk.  A run whose bucket started in this chunk
// is written to the bucket slot ("owned"); the first run of a chunk whose
// bucket started earlier goes to the chunk's carry slot for k_fixup.  Bucket
// key / Kj names the member whose points the entry indexes; the next entry's
// key is tracked ahead of the flush so its point can be prefetched from the
// right member.  Runs are stored unreduced as 29-bit slots, which the
// carry, fix-up and reduction kernels take as they are; only the window sums
// leave the 29-bit form (a conversion costs a product per component, too
// much for the divergent flush).
template <class F> struct Slot29 {
  static constexpr int NC = FieldIO<F>::NW / 8;  // base-field components per coordinate
  static constexpr int SW = 4 * NC * 9;          // words of one 29-bit XYZZ slot
};

// four 29-bit elements (36 words) as nine 16-byte stores
__device__ __forceinline__ void store4x29(uint32_t *d, const Fq29 &a, const Fq29 &b, const Fq29 &c, const Fq29 &e) {
  uint32_t w[36];
#pragma unroll
  for (int i = 0; i < 9; i++) {
    w[i] = a.l[i];
    w[9 + i] = b.l[i];
    w[18 + i] = c.l[i];
    w[27 + i] = e.l[i];
  }
  uint4 *d4 = reinterpret_cast<uint4 *>(d);
#pragma unroll
  for (int q = 0; q < 9; q++) d4[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}
__device__ __forceinline__ void store29(uint32_t *d, const XYZZ29 &a) { store4x29(d, a.X, a.Y, a.ZZ, a.ZZZ); }
// component order X.c0, X.c1, Y.c0, ... as the 2^256-word XYZZ<Fq2> slot
__device__ __forceinline__ void store29(uint32_t *d, const XYZZ2_29 &a) {
  store4x29(d, a.X.c0, a.X.c1, a.Y.c0, a.Y.c1);
  store4x29(d + 36, a.ZZ.c0, a.ZZ.c1, a.ZZZ.c0, a.ZZZ.c1);
}

__device__ __forceinline__ XYZZ29 inf29() { return {Fq29{}, f29_const(Q29::one), Fq29{}, Fq29{}}; }

template <class F>
__global__ void __launch_bounds__(256) k_accumulate29(const MsmSeg *__restrict__ segs, uint32_t Kj,
                                                      const uint32_t *__restrict__ sorted,
                                                      const uint32_t *__restrict__ off, uint32_t K, uint32_t L,
                                                      const uint32_t *__restrict__ pts29,
                                                      uint32_t *__restrict__ buckets29,
                                                      uint32_t *__restrict__ carry29,
                                                      uint32_t *__restrict__ carry_key) {
  static_assert(std::is_same<F, Fq>::value, "29-bit limbs are for the base field");
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t E = off[K];
  const uint64_t e0l = (uint64_t)t * L;
  if (e0l >= E) return;
  const uint32_t e0 = (uint32_t)e0l;
  const uint32_t e1 = min(e0 + L, E);
  carry_key[t] = 0xFFFFFFFFu;
  uint32_t key = find_key(off, K, e0);
  uint32_t kend = off[key + 1];
  bool owned = off[key] == e0;
  uint32_t mem = key / Kj;
  // table runs pass no pts29: each member reads its own table slice
  const uint32_t *points = pts29 ? pts29 + (uint64_t)segs[mem].qbase * 16 : segs[mem].points;
  uint32_t nkey = key, nkend = kend;  // bucket of the entry being prefetched
  XYZZ29 acc = inf29();
  bool ainf = true;
  IndexReader ir(sorted, e0);
  uint32_t nv = ir.at(e0);
  uint32_t nx[8], ny[8];
  {
    const uint32_t *pp = points + (uint64_t)(nv & 0x7FFFFFFFu) * 16;
    ld8(nx, pp);
    ld8(ny, pp + 8);
  }
  for (uint32_t e = e0; e < e1;) {
    const uint32_t v = nv;
    const Fq29 x = unpack29(nx);
    Fq29 y = unpack29(ny);
    e++;
    if (e < e1) {
      if (e == nkend) {
        nkey = key_from(off, K, nkey + 1, e);
        nkend = off[nkey + 1];
        const uint32_t m2 = nkey / Kj;
        if (m2 != mem) {
          mem = m2;
          points = pts29 ? pts29 + (uint64_t)segs[mem].qbase * 16 : segs[mem].points;
        }
      }
      nv = ir.at(e);
      const uint32_t *pp = points + (uint64_t)(nv & 0x7FFFFFFFu) * 16;
      ld8(nx, pp);
      ld8(ny, pp + 8);
    }
    const Fq29 yn = sub29<1>(Fq29{}, y);  // q - y for a negative digit
    if (v >> 31) y = yn;
    madd29(acc, ainf, x, y);
    if (e == kend || e == e1) {
      if (owned) {
        store29(buckets29 + (uint64_t)key * Slot29<F>::SW, acc);
      } else {
        store29(carry29 + (uint64_t)t * Slot29<F>::SW, acc);
        carry_key[t] = key;
      }
      acc = inf29();
      ainf = true;
      owned = true;
      key = nkey;
      kend = nkend;
    }
  }
}

__device__ __forceinline__ XYZZ2_29 inf2_29() { return {Fq2_29{}, f2one29(), Fq2_29{}, Fq2_29{}}; }

// signed entry v of a G2 run: the point in the 2^261 form, y negated for a
// negative digit
__device__ __forceinline__ void load2_29(const uint32_t *points, uint32_t v, Fq2_29 &x, Fq2_29 &y) {
  const uint32_t *pp = points + (uint64_t)(v & 0x7FFFFFFFu) * 32;
  uint32_t w[8];
  ld8(w, pp);
  x.c0 = unpack29(w);
  ld8(w, pp + 8);
  x.c1 = unpack29(w);
  ld8(w, pp + 16);
  y.c0 = unpack29(w);
  ld8(w, pp + 24);
  y.c1 = unpack29(w);
  if (v >> 31) {
    y.c0 = sub29<1>(Fq29{}, y.c0);
    y.c1 = sub29<1>(Fq29{}, y.c1);
  }
}

// G2 bucket loop on 29-bit limbs: as k_accumulate29, without the next-point
// prefetch (no registers to spare; the second resident wave hides the
// loads), held to two waves per SIMD.  The rare doubling reloads its point.
template <class F>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_accumulate29_g2(const MsmSeg *__restrict__ segs, uint32_t Kj, const uint32_t *__restrict__ sorted,
                  const uint32_t *__restrict__ off, uint32_t K, uint32_t L, const uint32_t *__restrict__ pts29,
                  uint32_t *__restrict__ buckets29, uint32_t *__restrict__ carry29,
                  uint32_t *__restrict__ carry_key) {
  static_assert(std::is_same<F, Fq2>::value, "the G2 loop");
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t E = off[K];
  const uint64_t e0l = (uint64_t)t * L;
  if (e0l >= E) return;
  const uint32_t e0 = (uint32_t)e0l;
  const uint32_t e1 = min(e0 + L, E);
  carry_key[t] = 0xFFFFFFFFu;
  uint32_t key = find_key(off, K, e0);
  uint32_t kend = off[key + 1];
  bool owned = off[key] == e0;
  uint32_t mem = key / Kj;
  const uint32_t *points = pts29 ? pts29 + (uint64_t)segs[mem].qbase * 32 : segs[mem].points;
  uint32_t nkey = key, nkend = kend;  // bucket of entry e after the increment
  XYZZ2_29 acc = inf2_29();
  bool ainf = true;
  IndexReader ir(sorted, e0);
  for (uint32_t e = e0; e < e1;) {
    const uint32_t v = ir.at(e);
    const uint32_t *cur = points;
    Fq2_29 x, y;
    load2_29(cur, v, x, y);
    e++;
    if (e < e1 && e == nkend) {
      nkey = key_from(off, K, nkey + 1, e);
      nkend = off[nkey + 1];
      const uint32_t m2 = nkey / Kj;
      if (m2 != mem) {
        mem = m2;
        points = pts29 ? pts29 + (uint64_t)segs[mem].qbase * 32 : segs[mem].points;
      }
    }
    if (!madd2_29(acc, ainf, x, y)) {
      Fq2_29 x2, y2;
      load2_29(cur, v, x2, y2);
      acc = dbl2_29_affine(x2, y2);
    }
    if (e == kend || e == e1) {
      if (owned) {
        store29(buckets29 + (uint64_t)key * Slot29<F>::SW, acc);
      } else {
        store29(carry29 + (uint64_t)t * Slot29<F>::SW, acc);
        carry_key[t] = key;
      }
      acc = inf2_29();
      ainf = true;
      owned = true;
      key = nkey;
      kend = nkend;
    }
  }
}

// 29-bit slot access for the carry, fix-up and reduction kernels
__device__ __forceinline__ void load4x29(const uint32_t *s, Fq29 &a, Fq29 &b, Fq29 &c, Fq29 &e) {
  const uint4 *s4 = reinterpret_cast<const uint4 *>(s);
  uint32_t w[36];
#pragma unroll
  for (int q = 0; q < 9; q++) {
    const uint4 v = s4[q];
    w[4 * q] = v.x;
    w[4 * q + 1] = v.y;
    w[4 * q + 2] = v.z;
    w[4 * q + 3] = v.w;
  }
#pragma unroll
  for (int i = 0; i < 9; i++) {
    a.l[i] = w[i];
    b.l[i] = w[9 + i];
    c.l[i] = w[18 + i];
    e.l[i] = w[27 + i];
  }
}
__device__ __forceinline__ void load29(const uint32_t *s, XYZZ29 &a) { load4x29(s, a.X, a.Y, a.ZZ, a.ZZZ); }
__device__ __forceinline__ void load29(const uint32_t *s, XYZZ2_29 &a) {
  load4x29(s, a.X.c0, a.X.c1, a.Y.c0, a.Y.c1);
  load4x29(s + 36, a.ZZ.c0, a.ZZ.c1, a.ZZZ.c0, a.ZZZ.c1);
}
template <class F> struct Acc29;
template <> struct Acc29<Fq> {
  using T = XYZZ29;
  __device__ __forceinline__ static T inf() { return inf29(); }
};
template <> struct Acc29<Fq2> {
  using T = XYZZ2_29;
  __device__ __forceinline__ static T inf() { return inf2_29(); }
};

// Carries of one bucket occupy the consecutive chunk slots t0+1 .. t1 of a
// bucket [s, e) (t0 = s/L, t1 = (e-1)/L).  A bucket far larger than a chunk
// (skewed digits: small witness values, a narrow top window) gives a long
// chain, so chains are summed by log-depth radix-8 rounds: in the round of
// step 8^r the slot at chain index i (i % 8^(r+1) == 0) adds the slots at
// i + k*8^r, k = 1..7, that lie in the chain.  The slots read are never
// written in the same round.  Slots are 29-bit (SW words).
template <class F>
__global__ void __launch_bounds__(256) k_carry_round(const uint32_t *__restrict__ off, uint32_t K, uint32_t L,
                                                     uint32_t *__restrict__ carry,
                                                     const uint32_t *__restrict__ carry_key, uint32_t step) {
  constexpr int SW = Slot29<F>::SW;
  using A = typename Acc29<F>::T;
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t E = off[K];
  if ((uint64_t)u * L >= E) return;
  const uint32_t key = carry_key[u];
  if (key == 0xFFFFFFFFu) return;
  const uint32_t t0 = off[key] / L, t1 = (off[key + 1] - 1) / L;
  const uint32_t i = u - (t0 + 1), len = t1 - t0;
  if (((uint64_t)i & (8ull * step - 1)) != 0 || (uint64_t)i + step >= len) return;
  A a;
  load29(carry + (uint64_t)u * SW, a);
#pragma unroll 1
  for (uint64_t k = 1; k < 8 && i + k * step < len; k++) {
    A b;
    load29(carry + (u + k * step) * SW, b);
    xadd29(a, b);
  }
  store29(carry + (uint64_t)u * SW, a);
}

// One thread per bucket: add the (reduced) chain head at slot t0+1.
template <class F>
__global__ void __launch_bounds__(256) k_fixup(const uint32_t *__restrict__ off, uint32_t K, uint32_t L,
                                               uint32_t *__restrict__ buckets,
                                               const uint32_t *__restrict__ carry) {
  constexpr int SW = Slot29<F>::SW;
  using A = typename Acc29<F>::T;
  uint32_t key = blockIdx.x * blockDim.x + threadIdx.x;
  if (key >= K) return;
  const uint32_t s = off[key], e = off[key + 1];
  if (e == s) return;
  const uint32_t t0 = s / L, t1 = (e - 1) / L;
  if (t1 == t0) return;
  A acc, c;
  load29(buckets + (uint64_t)key * SW, acc);
  load29(carry + (uint64_t)(t0 + 1) * SW, c);
  xadd29(acc, c);
  store29(buckets + (uint64_t)key * SW, acc);
}

// per (window, segment): sum_{b in seg} (b+1) * bucket_b
template <class F>
__global__ void __launch_bounds__(256) k_bucket_reduce(const uint32_t *__restrict__ off,
                                                       const uint32_t *__restrict__ buckets, uint32_t W,
                                                       uint32_t B, uint32_t Lr,
                                                       uint32_t *__restrict__ partial) {
  constexpr int SW = Slot29<F>::SW;
  using A = typename Acc29<F>::T;
  uint32_t S = B / Lr;
  uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= W * S) return;
  uint32_t w = tid / S, s = tid % S;
  uint32_t lo = s * Lr;
  A run = Acc29<F>::inf(), acc = Acc29<F>::inf();
#pragma unroll 1
  for (int b = (int)(lo + Lr) - 1; b >= (int)lo; b--) {
    uint32_t key = w * B + (uint32_t)b;
    if (off[key + 1] != off[key]) {
      A bk;
      load29(buckets + (uint64_t)key * SW, bk);
      xadd29(run, bk);
    }
    xadd29(acc, run);
  }
  xadd29(acc, xmul_small29(run, lo, Acc29<F>::inf()));
  store29(partial + (uint64_t)tid * SW, acc);
}

// one block per window (256 threads G1, 128 G2): sum of its S segment
// results, written as a 29-bit slot
template <class F>
__global__ void __launch_bounds__(256) k_window_reduce(const uint32_t *__restrict__ partial, uint32_t S,
                                                       uint32_t *__restrict__ wsum) {
  constexpr int SW = Slot29<F>::SW;
  using A = typename Acc29<F>::T;
  __shared__ uint32_t sh[256 * 36];  // blockDim * SW words
  uint32_t w = blockIdx.x;
  A acc = Acc29<F>::inf();
#pragma unroll 1
  for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
    A b;
    load29(partial + ((uint64_t)w * S + s) * SW, b);
    xadd29(acc, b);
  }
  store29(sh + threadIdx.x * SW, acc);
  __syncthreads();
  for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      A a, b;
      load29(sh + threadIdx.x * SW, a);
      load29(sh + (threadIdx.x + st) * SW, b);
      xadd29(a, b);
      store29(sh + threadIdx.x * SW, a);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int k = 0; k < SW; k++) wsum[(uint64_t)w * SW + k] = sh[k];
  }
}

// Window sums from nwin * S segment results (window-major).  One block per
// window leaves each thread S / blockDim serial additions; above one result
// per thread the sum runs in two launches: groups of blockDim results into
// tmp (nwin * S / blockDim slots, which may alias the bucket slots the
// segment results were made from), then one block per window over those.
template <class F>
void window_sums(const uint32_t *partial, uint32_t nwin, uint32_t S, uint32_t *tmp, uint32_t *wsum,
                 hipStream_t st) {
  const uint32_t TB = FieldIO<F>::NW == 16 ? 128 : 256;
  if (S > TB) {
    k_window_reduce<F><<<nwin * (S / TB), TB, 0, st>>>(partial, TB, tmp);
    k_window_reduce<F><<<nwin, TB, 0, st>>>(tmp, S / TB, wsum);
  } else {
    k_window_reduce<F><<<nwin, TB, 0, st>>>(partial, S, wsum);
  }
}

// One wave per member: Horner over its W window sums (c doublings each),
// written as a canonical affine point to the member's output.  The chain is
// serial, so the wave evaluates each formula cooperatively: the independent
// products of a step run on different lanes at once (coop29.h).  G1 keeps the
// point in every lane's registers as 29-bit limbs and exchanges the products
// through LDS; G2 keeps every value in 29-bit LDS slots, which holds its
// register count down.  Infinity is tracked in a flag every lane holds.
template <class F>
__global__ void __launch_bounds__(64) k_final(const MsmSeg *__restrict__ segs, const uint32_t *__restrict__ wsum,
                                              uint32_t W, uint32_t c) {
  const uint32_t j = blockIdx.x;
  uint32_t *out = segs[j].out;
  if (segs[j].n == 0) {
    if (threadIdx.x == 0) store_affine(out, Affine<F>{F::zero(), F::one()});
    return;
  }
  if constexpr (std::is_same<F, Fq>::value) {
    constexpr int SW = Slot29<F>::SW;
    __shared__ Fq29 sh[6];
    const uint32_t *ws = wsum + (uint64_t)j * W * SW;
    XYZZ29 acc;
    load29(ws + (uint64_t)(W - 1) * SW, acc);
    bool inf = is_zero29(acc.ZZ);
    for (int w = (int)W - 2; w >= 0; w--) {
      if (!inf)
        for (uint32_t i = 0; i < c; i++) xdbl29_coop(sh, acc);
      XYZZ29 q;
      load29(ws + (uint64_t)w * SW, q);
      if (is_zero29(q.ZZ)) continue;
      if (inf) {
        acc = q;
        inf = false;
        continue;
      }
      xadd29_coop(sh, acc, q, inf);
    }
    if (threadIdx.x == 0) {
      XYZZ<F> a;
      from29_words(acc.X, a.X.v);
      from29_words(acc.Y, a.Y.v);
      from29_words(acc.ZZ, a.ZZ.v);
      from29_words(acc.ZZZ, a.ZZZ.v);
      if (inf) a = XYZZ<F>::inf();
      store_affine(out, xyzz_to_affine(a));
    }
  } else {
    using C = Coop29<F>;
    constexpr int NC = C::NC, SW = NC * 9;
    __shared__ uint32_t sh[C::kSlots * 9];
    __shared__ C29Op prog[C::kOps];
    C::load(prog);
    const uint32_t *ws = wsum + (uint64_t)j * W * SW;
    // acc = window W-1; the zero slot follows the addend
    for (uint32_t i = threadIdx.x; i < (uint32_t)SW; i += blockDim.x) {
      sh[i] = ws[(uint64_t)(W - 1) * SW + i];
      sh[2 * SW + i % 9] = 0;
    }
    __syncthreads();
    bool inf = c29_is_zero(sh, C::ZZ);
    for (int w = (int)W - 2; w >= 0; w--) {
      if (!inf)
        for (uint32_t i = 0; i < c; i++) C::dbl(sh, prog);
      for (uint32_t i = threadIdx.x; i < (uint32_t)SW; i += blockDim.x) sh[SW + i] = ws[(uint64_t)w * SW + i];
      __syncthreads();
      if (c29_is_zero(sh, C::QZZ)) continue;
      if (inf) {
        for (uint32_t i = threadIdx.x; i < (uint32_t)SW; i += blockDim.x) sh[i] = sh[SW + i];
        __syncthreads();
        inf = false;
        continue;
      }
      C::add1(sh, prog);
      if (c29_is_zero(sh, C::PP)) {
        if (c29_is_zero(sh, C::RR)) C::dbl(sh, prog);  // acc == q
        else inf = true;                                // acc == -q
        continue;
      }
      C::add2(sh, prog);
    }
    if (threadIdx.x == 0) {
      uint32_t wds[4 * FieldIO<F>::NW];
      for (int k = 0; k < NC; k++) {
        uint32_t o[8];
        from29_words(c29_ld(sh, k), o);
        for (int i = 0; i < 8; i++) wds[8 * k + i] = o[i];
      }
      XYZZ<F> acc = load_xyzz<F>(wds);
      if (inf) acc = XYZZ<F>::inf();
      store_affine(out, xyzz_to_affine(acc));
    }
  }
}

// The sort's placement passes over the counted and scanned digits: pass 1
// (bins by hi) and pass 2 (buckets by lo), for a group of J members of pl.W
// windows whose scalars have Wwin windows each (pl.W, or Wt for the one-window
// members of table runs), into offsets (Ktot + 1) and sorted.
static inline void sort_place(const MsmPlan &pl, uint32_t Wwin, const uint32_t *digits, const MsmSeg *segs,
                              const uint32_t *cnt, uint32_t *tix, uint16_t *tlo, uint32_t *offsets,
                              uint32_t *sorted, hipStream_t st) {
  const uint32_t sb = pl.c - 1 - pl.hb, HB = 1u << pl.hb;
  const uint32_t nblk = (uint32_t)(pl.W * pl.tiles), Wtot = pl.J * pl.W, Ktot = pl.J * pl.K;
  if (pl.max_entries < (1ull << 27))
    k_sort_scatter1<true><<<nblk, SORT_THREADS, 0, st>>>(digits, segs, pl.J, sb, pl.st, HB, pl.W, Wwin, cnt, tix,
                                                          tlo, sorted);
  else
    k_sort_scatter1<false><<<nblk, SORT_THREADS, 0, st>>>(digits, segs, pl.J, sb, pl.st, HB, pl.W, Wwin, cnt, tix,
                                                           tlo, sorted);
  // the 8192-entry tile keeps 49 KB of staging in LDS beside 3 * 2^sb counter
  // words: used up to sb = 10 (61 KB); wider windows take the 4096 tile
  if (pl.max_entries < (1ull << 27) || sb > 10)
    k_sort_pass2<SORT_TILE><<<Wtot * HB, SORT_THREADS, 3 * (1u << sb) * 4, st>>>(
        tix, tlo, cnt, segs, HB, sb, pl.st, pl.W, Wwin, Ktot, offsets, sorted);
  else
    k_sort_pass2<2 * SORT_TILE><<<Wtot * HB, SORT_THREADS, 3 * (1u << sb) * 4, st>>>(
        tix, tlo, cnt, segs, HB, sb, pl.st, pl.W, Wwin, Ktot, offsets, sorted);
}

// Enqueue the whole pipeline for one group on stream st.  segs is the
// group's device descriptor table (J + 1 entries); ws is carved in the order
// msm_workspace_bytes assumes.  No allocation, no host synchronisation and no
// memset nodes, so the sequence can be captured into a graph.
template <class F>
int msm_run(void *ws, const MsmPlan &pl, const MsmSeg *segs, hipStream_t st, StageTimer *tm) {
  const bool g2 = FieldIO<F>::NW == 16;
  Carver cv(ws);
  const uint32_t hi_bits = pl.hb;
  const uint32_t sb = pl.c - 1 - hi_bits, HB = 1u << hi_bits;
  const uint32_t nblk = (uint32_t)(pl.W * pl.tiles);
  const uint32_t ncnt = nblk * HB;
  const uint32_t nb = (ncnt + SCAN_TILE - 1) / SCAN_TILE;
  const uint32_t Ktot = pl.J * pl.K, Wtot = pl.J * pl.W;
  uint32_t *digits = cv.take<uint32_t>(pl.max_entries);
  uint32_t *sorted = cv.take<uint32_t>(pl.max_entries + 4);  // + 4: quad loads
  constexpr int PW = 2 * FieldIO<F>::NW, SW = Slot29<F>::SW;
  char *sortbuf = cv.take<char>(msm_sortbuf_bytes(pl, g2));
  Carver cs(sortbuf);
  uint32_t *pts29 = cs.take<uint32_t>((size_t)pl.n * PW);  // written by k_digits
  char *after = sortbuf + Carver::round((size_t)pl.n * PW * 4);
  Carver cs1(after);
  uint32_t *tix = cs1.take<uint32_t>(pl.max_entries);
  uint16_t *tlo = cs1.take<uint16_t>(pl.max_entries);
  uint32_t *cnt = cv.take<uint32_t>((size_t)ncnt + 1);
  uint32_t *bsum = cv.take<uint32_t>(nb);
  uint32_t *offsets = cv.take<uint32_t>((size_t)Ktot + 1);
  uint32_t *carry_key = cv.take<uint32_t>(pl.max_chunks);
  uint32_t *partial = cv.take<uint32_t>((size_t)Wtot * pl.S * SW);
  uint32_t *wsum = cv.take<uint32_t>((size_t)Wtot * SW);
  // the pass-1 arrays are free once the sort is done: bucket and carry slots
  Carver c29(after);
  uint32_t *b29 = c29.take<uint32_t>((size_t)Ktot * SW);
  uint32_t *cr29 = c29.take<uint32_t>((size_t)pl.max_chunks * SW);

  if (tm) tm->mark(0, st);
  if (pl.n > 0) {
    k_digits<F><<<(uint32_t)pl.tiles, 256, pl.W * HB * 4, st>>>(segs, pl.J, pl.c, pl.W, sb, pl.st, HB, digits, cnt,
                                                                   pts29);
    CSK_CHECK_LAUNCH();
    k_scan_reduce<<<nb, SCAN_THREADS, 0, st>>>(cnt, ncnt, bsum);
    k_scan_blocks<<<1, 1024, 0, st>>>(bsum, nb);
    k_scan_final<<<nb, SCAN_THREADS, 0, st>>>(cnt, ncnt, bsum, cnt);
    sort_place(pl, pl.W, digits, segs, cnt, tix, tlo, offsets, sorted, st);
    CSK_CHECK_LAUNCH();
    const uint32_t TB = g2 ? 128 : 256;
    uint32_t ga = (uint32_t)((pl.max_chunks + TB - 1) / TB);
    if (tm) tm->mark(1, st);
    if constexpr (std::is_same<F, Fq>::value)
      k_accumulate29<F><<<ga, TB, 0, st>>>(segs, pl.K, sorted, offsets, Ktot, pl.L, pts29, b29, cr29, carry_key);
    else
      k_accumulate29_g2<F><<<ga, TB, 0, st>>>(segs, pl.K, sorted, offsets, Ktot, pl.L, pts29, b29, cr29, carry_key);
    CSK_CHECK_LAUNCH();
    if (tm) tm->mark(2, st);
    // chains can be at most max_chunks long
    for (uint64_t step = 1; step < pl.max_chunks; step *= 8)
      k_carry_round<F><<<ga, TB, 0, st>>>(offsets, Ktot, pl.L, cr29, carry_key, (uint32_t)step);
    CSK_CHECK_LAUNCH();
    k_fixup<F><<<(Ktot + TB - 1) / TB, TB, 0, st>>>(offsets, Ktot, pl.L, b29, cr29);
    CSK_CHECK_LAUNCH();
    uint32_t nseg = Wtot * pl.S;
    k_bucket_reduce<F><<<(nseg + TB - 1) / TB, TB, 0, st>>>(offsets, b29, Wtot, pl.B, pl.Lr, partial);
    CSK_CHECK_LAUNCH();
    window_sums<F>(partial, Wtot, pl.S, b29, wsum, st);  // the bucket slots are free again
    CSK_CHECK_LAUNCH();
  }
  k_final<F><<<pl.J, 64, 0, st>>>(segs, wsum, pl.W, pl.c);
  CSK_CHECK_LAUNCH();
  if (tm) tm->mark(3, st);
  return CSK_OK;
}


// ------------------------------------------------------------ fixed-base tables
// A table of points P_0 .. P_(n-1) for window width c holds Q[w * n_pad + i] =
// 2^(c w) P_i for w < Wt = floor(254 / c) + 1 (n_pad = n rounded up to the
// sort tile), in the 2^261 form the bucket loops read, and a bit per point,
// 1 unless P_i is at infinity.  With the signed c-bit digits d_(i,w) of k_i,
// sum_i k_i P_i = sum_(w, i) d_(i,w) Q[w * n_pad + i]: the multiexp runs as a
// group of Wt one-window members (member w: digit w of every scalar, points
// Q[w * n_pad ..]), which sorts and accumulates exactly as a plain group;
// then bucket b of the sum is the sum over w of member w's bucket b, reduced
// once, and no doublings are needed between windows.

// one thread per point: its Wt table entries (an inversion each) and its bit;
// grid covers n_pad (a multiple of the sort tile) exactly
template <class F>
__global__ void __launch_bounds__(256) k_table_build(const uint32_t *__restrict__ pts, uint64_t n, uint64_t n_pad,
                                                     uint32_t c, uint32_t Wt, uint32_t *__restrict__ table,
                                                     uint32_t *__restrict__ live) {
  constexpr int PW = 2 * FieldIO<F>::NW;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  Affine<F> p = {F::zero(), F::zero()};
  bool inf = true;
  if (i < n) {
    p = load_affine<F>(pts + i * PW);
    inf = affine_is_inf(p);
  }
  // 32 points per word; i & 31 == threadIdx.x & 31 (blocks of 256)
  const uint64_t m = __ballot(!inf);
  if ((threadIdx.x & 31) == 0 && i < n_pad) live[i >> 5] = (uint32_t)(m >> (threadIdx.x & 32));
  if (inf) return;
  XYZZ<F> acc = XYZZ<F>::from_affine(p);
  store_point29(table + i * PW, p);
  for (uint32_t w = 1; w < Wt; w++) {
    for (uint32_t k = 0; k < c; k++) acc = xyzz_dbl(acc);
    store_point29(table + ((uint64_t)w * n_pad + i) * PW, xyzz_to_affine(acc));
  }
}

namespace {  // curve-independent table kernels: one private copy per translation unit

// Digits of the table group: one workgroup per sort tile of scalars and run
// p (blockIdx.y; tiles past the run's table end return).  Member m = p * Wt +
// w (window w of run p) has digit array D[pbase_m + i], pbase_m = pbase_(p*Wt)
// + w * n_pad (as k_digits writes a one-window member at its concatenated
// points), empty (0xFFFFFFFF) for a zero digit, a point at infinity or i >= n.
// Counts go to cnt[cbase_m + hi * NT + tile], cbase_m = cbase_(p*Wt) + w * HB
// * NT, the member's pass-1 counters.
__global__ void __launch_bounds__(256) k_digits_tab(TabArgs args, const MsmSeg *__restrict__ segs, uint32_t c,
                                                    uint32_t Wt, uint32_t s, uint32_t ts, uint32_t HB,
                                                    uint32_t *__restrict__ digits, uint32_t *__restrict__ cnt) {
  extern __shared__ uint32_t hist[];  // Wt * HB counters
  const uint32_t tile = blockIdx.x, p = blockIdx.y;
  const MsmSeg &g0 = segs[p * Wt];
  const uint32_t NT = g0.NT;
  if (tile >= NT) return;  // the whole workgroup: this run's table has fewer tiles
  const uint64_t n = args.n[p], n_pad = g0.n;
  const uint32_t *__restrict__ scalars = args.s[p];
  const uint32_t *__restrict__ live = args.live[p];
  digits += g0.pbase;
  cnt += g0.cbase;
  for (uint32_t i = threadIdx.x; i < Wt * HB; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const uint32_t B = 1u << (c - 1);
  const uint32_t mask = (1u << c) - 1;
  for (uint32_t t = threadIdx.x; t < kSortTile; t += blockDim.x) {
    const uint64_t i = (uint64_t)tile * kSortTile + t;
    if (i >= n) {
      for (uint32_t w = 0; w < Wt; w++) digits[(uint64_t)w * n_pad + i] = 0xFFFFFFFFu;
      continue;
    }
    uint32_t k[9];
    ld8(k, scalars + i * 8);
    k[8] = 0;
    reduce_mod_r(k);
    const bool zero = ((live[i >> 5] >> (i & 31)) & 1u) == 0;
    uint32_t carry = 0;
    for (uint32_t w = 0; w < Wt; w++) {
      uint32_t off = w * c;
      uint32_t li = off >> 5, sh = off & 31;
      uint64_t two = (uint64_t)k[li] | ((uint64_t)(li < 7 ? k[li + 1] : 0u) << 32);
      uint32_t raw = ((uint32_t)(two >> sh) & mask) + carry;
      uint32_t mag, neg;
      if (raw > B) {
        mag = (1u << c) - raw;
        neg = 1;
        carry = 1;
      } else {
        mag = raw;
        neg = 0;
        carry = 0;
      }
      const bool lv = mag != 0 && !zero;
      digits[(uint64_t)w * n_pad + i] = lv ? ((mag - 1) | (neg << 31)) : 0xFFFFFFFFu;
      if (lv) atomicAdd(&hist[w * HB + ((mag - 1) >> (w == Wt - 1 ? ts : s))], 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < Wt * HB; i += blockDim.x) cnt[(uint64_t)i * NT + tile] = hist[i];
}

// merged[p * B + b] = sum over the J members of vector p (blockIdx.y) of their
// bucket b (empty buckets skipped)
template <class F>
__global__ void __launch_bounds__(256) k_merge_members(const uint32_t *__restrict__ off,
                                                       const uint32_t *__restrict__ buckets, uint32_t J, uint32_t B,
                                                       uint32_t *__restrict__ merged) {
  constexpr int SW = Slot29<F>::SW;
  using A = typename Acc29<F>::T;
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const uint64_t p = blockIdx.y;
  off += p * J * B;
  buckets += p * J * B * SW;
  merged += p * B * SW;
  A acc = Acc29<F>::inf();
#pragma unroll 1
  for (uint32_t j = 0; j < J; j++) {
    const uint32_t key = j * B + b;
    if (off[key + 1] != off[key]) {
      A bk;
      load29(buckets + (uint64_t)key * SW, bk);
      xadd29(acc, bk);
    }
  }
  store29(merged + (uint64_t)b * SW, acc);
}

// o[k] = k for k <= B: bucket offsets under which every merged bucket is read
__global__ void __launch_bounds__(256) k_iota(uint32_t *__restrict__ o, uint32_t count) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) o[k] = k;
}

// one output descriptor per scalar vector (what k_final reads), and a sentinel
__global__ void k_put_outs(TabArgs args, uint32_t count, MsmSeg *osegs) {
  const uint32_t p = threadIdx.x;
  if (p > count) return;
  MsmSeg g = {};
  if (p < count) {
    g.out = args.o[p];
    g.n = 1;
  }
  osegs[p] = g;
}

}  // namespace

// Workspace of a table run over count scalar vectors, carved in this order
// by msm_run_table.
inline size_t msm_table_workspace_bytes(const MsmPlan &pl, bool g2, uint32_t Sm, uint32_t count) {
  const uint32_t hi_bits = pl.hb;
  const uint64_t ncnt = (uint64_t)pl.W * (1u << hi_bits) * pl.tiles;
  const uint64_t nb = (ncnt + SCAN_TILE - 1) / SCAN_TILE;
  const uint64_t Ktot = (uint64_t)pl.J * pl.K;
  const uint64_t sb = g2 ? 288 : 144;
  const uint64_t sort = Carver::round(pl.max_entries * 4) + pl.max_entries * 2;
  const uint64_t slots = Carver::round(Ktot * sb) + pl.max_chunks * sb;
  size_t b = 0;
  b += Carver::round(pl.max_entries * 4);        // digits
  b += Carver::round((pl.max_entries + 4) * 4);  // sorted
  b += Carver::round(sort > slots ? sort : slots);
  b += Carver::round((ncnt + 1) * 4);
  b += Carver::round(nb * 4);
  b += Carver::round((Ktot + 1) * 4);
  b += Carver::round(pl.max_chunks * 4);
  b += Carver::round((uint64_t)count * pl.B * sb);        // merged buckets
  b += Carver::round(((uint64_t)count * pl.B + 1) * 4);   // iota offsets
  b += Carver::round((uint64_t)count * Sm * sb);          // partials of the merged windows
  b += Carver::round((uint64_t)count * sb);               // their sums
  b += Carver::round((uint64_t)(count + 1) * sizeof(MsmSeg));  // output descriptors
  b += Carver::round(((uint64_t)pl.J + 1) * sizeof(MsmSeg));   // member descriptors
  return b + 4096;
}

// Enqueue count table multiexps (run p: scalar vector args.s[p] over its
// table -> args.o[p]): pl is the group plan of the count * Wt one-window
// members m = p * Wt + w (J = count * Wt, W = 1), hsegs their host
// descriptors (member m reads its table's slice of window w through its points
// pointer), Lrm / Sm the reduction split of a merged window.
template <class F>
int msm_run_table(void *ws, const MsmPlan &pl, const MsmSeg *hsegs, const TabArgs &args, uint32_t count,
                  uint32_t max_nt, uint32_t Lrm, uint32_t Sm, hipStream_t st) {
  const bool g2 = FieldIO<F>::NW == 16;
  constexpr int SW = Slot29<F>::SW;
  Carver cv(ws);
  const uint32_t hi_bits = pl.hb;
  const uint32_t sb = pl.c - 1 - hi_bits, HB = 1u << hi_bits;
  const uint32_t nblk = (uint32_t)(pl.W * pl.tiles);
  const uint32_t ncnt = nblk * HB;
  const uint32_t nb = (ncnt + SCAN_TILE - 1) / SCAN_TILE;
  const uint32_t Ktot = pl.J * pl.K, Wtot = pl.J * pl.W;
  uint32_t *digits = cv.take<uint32_t>(pl.max_entries);
  uint32_t *sorted = cv.take<uint32_t>(pl.max_entries + 4);
  const uint64_t sortb = Carver::round(pl.max_entries * 4) + pl.max_entries * 2;
  const uint64_t slotb = Carver::round((uint64_t)Ktot * SW * 4) + pl.max_chunks * SW * 4;
  char *sortbuf = cv.take<char>(sortb > slotb ? sortb : slotb);
  Carver cs1(sortbuf);
  uint32_t *tix = cs1.take<uint32_t>(pl.max_entries);
  uint16_t *tlo = cs1.take<uint16_t>(pl.max_entries);
  Carver c29(sortbuf);  // the slots reuse the pass-1 arrays once the sort is done
  uint32_t *b29 = c29.take<uint32_t>((size_t)Ktot * SW);
  uint32_t *cr29 = c29.take<uint32_t>((size_t)pl.max_chunks * SW);
  uint32_t *cnt = cv.take<uint32_t>((size_t)ncnt + 1);
  uint32_t *bsum = cv.take<uint32_t>(nb);
  uint32_t *offsets = cv.take<uint32_t>((size_t)Ktot + 1);
  uint32_t *carry_key = cv.take<uint32_t>(pl.max_chunks);
  uint32_t *merged = cv.take<uint32_t>((size_t)count * pl.B * SW);
  uint32_t *iota = cv.take<uint32_t>((size_t)count * pl.B + 1);
  uint32_t *partial = cv.take<uint32_t>((size_t)count * Sm * SW);
  uint32_t *wsum = cv.take<uint32_t>((size_t)count * SW);
  MsmSeg *osegs = cv.take<MsmSeg>(count + 1);
  MsmSeg *segs = cv.take<MsmSeg>(pl.J + 1);
  const uint32_t Wt = pl.J / count;

  CSK_HIP(hipMemcpyAsync(segs, hsegs, ((size_t)pl.J + 1) * sizeof(MsmSeg), hipMemcpyHostToDevice, st));
  k_put_outs<<<1, 128, 0, st>>>(args, count, osegs);
  k_digits_tab<<<dim3(max_nt, count), 256, Wt * HB * 4, st>>>(args, segs, pl.c, Wt, sb, pl.st, HB, digits, cnt);
  CSK_CHECK_LAUNCH();
  k_scan_reduce<<<nb, SCAN_THREADS, 0, st>>>(cnt, ncnt, bsum);
  k_scan_blocks<<<1, 1024, 0, st>>>(bsum, nb);
  k_scan_final<<<nb, SCAN_THREADS, 0, st>>>(cnt, ncnt, bsum, cnt);
  sort_place(pl, Wt, digits, segs, cnt, tix, tlo, offsets, sorted, st);
  CSK_CHECK_LAUNCH();
  const uint32_t TB = g2 ? 128 : 256;
  const uint32_t ga = (uint32_t)((pl.max_chunks + TB - 1) / TB);
  if constexpr (std::is_same<F, Fq>::value)
    k_accumulate29<F><<<ga, TB, 0, st>>>(segs, pl.K, sorted, offsets, Ktot, pl.L, nullptr, b29, cr29, carry_key);
  else
    k_accumulate29_g2<F><<<ga, TB, 0, st>>>(segs, pl.K, sorted, offsets, Ktot, pl.L, nullptr, b29, cr29,
                                             carry_key);
  CSK_CHECK_LAUNCH();
  for (uint64_t step = 1; step < pl.max_chunks; step *= 8)
    k_carry_round<F><<<ga, TB, 0, st>>>(offsets, Ktot, pl.L, cr29, carry_key, (uint32_t)step);
  CSK_CHECK_LAUNCH();
  k_fixup<F><<<(Ktot + TB - 1) / TB, TB, 0, st>>>(offsets, Ktot, pl.L, b29, cr29);
  k_merge_members<F><<<dim3((pl.B + TB - 1) / TB, count), TB, 0, st>>>(offsets, b29, Wt, pl.B, merged);
  k_iota<<<(uint32_t)(((uint64_t)count * pl.B + 1 + 255) / 256), 256, 0, st>>>(iota, count * pl.B + 1);
  CSK_CHECK_LAUNCH();
  k_bucket_reduce<F><<<(count * Sm + TB - 1) / TB, TB, 0, st>>>(iota, merged, count, pl.B, Lrm, partial);
  window_sums<F>(partial, count, Sm, merged, wsum, st);  // the merged buckets are free again
  k_final<F><<<count, 64, 0, st>>>(osegs, wsum, 1, pl.c);
  CSK_CHECK_LAUNCH();
  return CSK_OK;
}

template <class F>
int msm_table_build(const uint32_t *pts, uint64_t n, uint64_t n_pad, uint32_t c, uint32_t Wt, uint32_t *table,
                    uint32_t *live, hipStream_t st) {
  k_table_build<F><<<(uint32_t)(n_pad / 256), 256, 0, st>>>(pts, n, n_pad, c, Wt, table, live);
  CSK_CHECK_LAUNCH();
  return CSK_OK;
}

}  // namespace csk
