// Base-field (and Fq2) arithmetic on 29-bit limbs for the G1 and G2 bucket
// loops (gfx950).
//
// An element is 9 limbs of 29 bits (261 bits) in Montgomery form with radix
// R' = 2^261.  With 29-bit limbs a limb product is below 2^58, so a product-
// scanning column of up to 9 + 9 products plus the carry-in stays below
// 2^64: every limb product is one v_mad_u64_u32 into a 64-bit column sum
// with no carry word (the 32-bit-limb product needs a v_addc per product).
// That is 162 multiply-adds and no carry instructions per product against
// 128 multiply-adds and 128 carry adds on 32-bit limbs.
//
// Ranges.  R' exceeds 170q, so for inputs A, B with A*B < q*R' (e.g. both
// below 13q) the Montgomery product is below 2q.  Sums and differences are
// therefore left unreduced: a - b is computed as a + Kq - b for a multiple
// Kq above b's range, with limbs normalised by one carry pass.  The
// G1 loop keeps X < 8q, Y, ZZ, ZZZ < 2q (see madd29).
//
// Conversion.  Points enter as the library's 8 x u32 words in Montgomery
// form for R = 2^256; to29 multiplies by 2^266 mod q (x 2^256 x 2^266 /
// 2^261 = x 2^261), stored packed back into 8 words (values are < q).
// from29 multiplies by 2^256 mod q, which returns to R = 2^256 form.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace csk {

struct Fq29 {
  uint32_t l[9];
};

struct Q29 {
  static constexpr uint32_t M = (1u << 29) - 1;
  static constexpr uint32_t np = 75916169u;  // -q^-1 mod 2^29
  static constexpr uint32_t p[9] = {410844487u, 17064118u, 477274959u, 47522512u, 361093496u,
                                    47923392u,  10936641u, 240920116u, 3171406u};
  static constexpr uint32_t one[9] = {360500257u, 337389400u, 408039635u, 21759001u, 178483129u,
                                      490881230u, 299191303u, 86689704u,  903222u};  // 2^261 mod q
  static constexpr uint32_t to[9] = {322215073u, 442336424u, 171859116u, 268585440u, 314135016u,
                                     244503300u, 348886451u, 68918589u,  360451u};  // 2^266 mod q
  static constexpr uint32_t from[9] = {93261213u,  451550318u, 297979764u, 299258347u, 342016167u,
                                       297253948u, 482187706u, 406012155u, 920183u};  // 2^256 mod q
  static constexpr uint32_t p4[9] = {32765212u,  68256475u, 298487100u, 190090051u, 370632160u,
                                     191693570u, 43746564u, 426809552u, 12685625u};  // 4q
};

// K*q with every limb but the top raised by 2^29 - 1 (2^29 for limb 0) and
// borrowed back from the next: each of the low 8 limbs is >= 2^29 - 1, so
// a_i + d_i - b_i is never negative for a normalised b.
template <int K> struct QMul;
template <> struct QMul<1> {
  static constexpr uint32_t d[9] = {947715399u, 553935029u, 1014145870u, 584393423u, 897964407u,
                                    584794303u, 547807552u, 777791027u,  3171405u};
};
template <> struct QMul<2> {
  static constexpr uint32_t d[9] = {821688974u, 570999148u, 954549917u,  631915936u, 722186991u,
                                    632717696u, 558744193u, 1018711143u, 6342811u};
};
template <> struct QMul<4> {
  static constexpr uint32_t d[9] = {569636124u, 605127386u, 835358011u, 726960962u, 907503071u,
                                    728564481u, 580617475u, 963680463u, 12685624u};
};
template <> struct QMul<6> {
  static constexpr uint32_t d[9] = {854454186u, 639255623u, 716166105u, 822005988u, 555948239u,
                                    824411267u, 602490757u, 908649783u, 19028437u};
};
template <> struct QMul<8> {
  static constexpr uint32_t d[9] = {602401336u, 673383861u, 596974199u, 917051014u, 741264319u,
                                    920258052u, 624364039u, 853619103u, 25371250u};
};

__device__ __forceinline__ Fq29 f29_const(const uint32_t (&c)[9]) {
  Fq29 r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.l[i] = c[i];
  return r;
}

// acc += x*y, one v_mad_u64_u32.  The empty asm makes each partial sum
// opaque, so a column stays one serial chain that starts from the carry-in
// (left to itself the compiler sums a column from zero and adds the carry-in
// with an extra 64-bit add).  Column sums stay below 2^64.
__device__ __forceinline__ void mad29(uint64_t &acc, uint32_t x, uint32_t y) {
  acc += (uint64_t)x * y;
  asm("" : "+v"(acc));
}
__device__ __forceinline__ void mad29_s(uint64_t &acc, uint32_t x, uint32_t y) { mad29(acc, x, y); }

// Montgomery product a*b*2^-261 by product scanning.  Needs the column sums
// below 2^64: both operands' limbs below 2^30, or one operand normalised
// (limbs < 2^29) and the other's limbs below 2^31.  Output normalised, < 2q
// when a*b < q*2^261.
__device__ __forceinline__ Fq29 mul29(const Fq29 &a, const Fq29 &b) {
  uint32_t m[9];
  Fq29 r;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < 17; k++) {
#pragma unroll
    for (int j = (k > 8 ? k - 8 : 0); j <= (k < 8 ? k : 8); j++) mad29(acc, a.l[j], b.l[k - j]);
#pragma unroll
    for (int j = (k > 8 ? k - 8 : 0); j < (k < 9 ? k : 9); j++) mad29_s(acc, m[j], Q29::p[k - j]);
    if (k < 9) {
      m[k] = ((uint32_t)acc * Q29::np) & Q29::M;
      mad29_s(acc, m[k], Q29::p[0]);
    } else {
      r.l[k - 9] = (uint32_t)acc & Q29::M;
    }
    acc >>= 29;
  }
  r.l[8] = (uint32_t)acc;
  return r;
}

// Two independent products, their column chains interleaved: back-to-back
// dependent v_mad_u64_u32 need a wait state on gfx950, alternating chains
// fill it with useful work.
__device__ __forceinline__ void mul29x2(const Fq29 &a, const Fq29 &b, const Fq29 &c, const Fq29 &d, Fq29 &ab,
                                        Fq29 &cd) {
  uint32_t m[9], n[9];
  uint64_t acc = 0, acd = 0;
#pragma unroll
  for (int k = 0; k < 17; k++) {
#pragma unroll
    for (int j = (k > 8 ? k - 8 : 0); j <= (k < 8 ? k : 8); j++) {
      mad29(acc, a.l[j], b.l[k - j]);
      mad29(acd, c.l[j], d.l[k - j]);
    }
#pragma unroll
    for (int j = (k > 8 ? k - 8 : 0); j < (k < 9 ? k : 9); j++) {
      mad29(acc, m[j], Q29::p[k - j]);
      mad29(acd, n[j], Q29::p[k - j]);
    }
    if (k < 9) {
      m[k] = ((uint32_t)acc * Q29::np) & Q29::M;
      n[k] = ((uint32_t)acd * Q29::np) & Q29::M;
      mad29(acc, m[k], Q29::p[0]);
      mad29(acd, n[k], Q29::p[0]);
    } else {
      ab.l[k - 9] = (uint32_t)acc & Q29::M;
      cd.l[k - 9] = (uint32_t)acd & Q29::M;
    }
    acc >>= 29;
    acd >>= 29;
  }
  ab.l[8] = (uint32_t)acc;
  cd.l[8] = (uint32_t)acd;
}

// a^2 and c^2 (normalised a, c), interleaved as mul29x2: each cross product
// a_i a_j (i < j) once, on the doubled limb 2 a_i (< 2^30), 126 multiply-adds
// per square instead of 162.
__device__ __forceinline__ void sqr29x2(const Fq29 &a, const Fq29 &c, Fq29 &aa, Fq29 &cc) {
  uint32_t a2[9], c2[9], m[9], n[9];
#pragma unroll
  for (int i = 0; i < 9; i++) {
    a2[i] = a.l[i] << 1;
    c2[i] = c.l[i] << 1;
  }
  uint64_t acc = 0, acd = 0;
#pragma unroll
  for (int k = 0; k < 17; k++) {
#pragma unroll
    for (int j = (k > 8 ? k - 8 : 0); 2 * j < k; j++) {
      mad29(acc, a2[j], a.l[k - j]);
      mad29(acd, c2[j], c.l[k - j]);
    }
    if ((k & 1) == 0) {
      mad29(acc, a.l[k / 2], a.l[k / 2]);
      mad29(acd, c.l[k / 2], c.l[k / 2]);
    }
#pragma unroll
    for (int j = (k > 8 ? k - 8 : 0); j < (k < 9 ? k : 9); j++) {
      mad29(acc, m[j], Q29::p[k - j]);
      mad29(acd, n[j], Q29::p[k - j]);
    }
    if (k < 9) {
      m[k] = ((uint32_t)acc * Q29::np) & Q29::M;
      n[k] = ((uint32_t)acd * Q29::np) & Q29::M;
      mad29(acc, m[k], Q29::p[0]);
      mad29(acd, n[k], Q29::p[0]);
    } else {
      aa.l[k - 9] = (uint32_t)acc & Q29::M;
      cc.l[k - 9] = (uint32_t)acd & Q29::M;
    }
    acc >>= 29;
    acd >>= 29;
  }
  aa.l[8] = (uint32_t)acc;
  cc.l[8] = (uint32_t)acd;
}

// Montgomery reduction of a*b + c*d in one pass (normalised operands): the
// a*b products run on a second chain merged into the column once per
// column, the c*d and reduction products on the main chain.  Output < 2q
// when a*b + c*d < q*2^261.
__device__ __forceinline__ Fq29 dot29(const Fq29 &a, const Fq29 &b, const Fq29 &c, const Fq29 &d) {
  uint32_t m[9];
  Fq29 r;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < 17; k++) {
    uint64_t u = 0;
#pragma unroll
    for (int j = (k > 8 ? k - 8 : 0); j <= (k < 8 ? k : 8); j++) {
      mad29(u, a.l[j], b.l[k - j]);
      mad29(acc, c.l[j], d.l[k - j]);
    }
#pragma unroll
    for (int j = (k > 8 ? k - 8 : 0); j < (k < 9 ? k : 9); j++) mad29(acc, m[j], Q29::p[k - j]);
    acc += u;
    if (k < 9) {
      m[k] = ((uint32_t)acc * Q29::np) & Q29::M;
      mad29(acc, m[k], Q29::p[0]);
    } else {
      r.l[k - 9] = (uint32_t)acc & Q29::M;
    }
    acc >>= 29;
  }
  r.l[8] = (uint32_t)acc;
  return r;
}

// Latency-ordered forms of mul29 / dot29 for code that runs one wave at a
// time (the window combination): the limb products go row by row into 17
// independent column sums, so consecutive multiply-adds never wait for each
// other, and only the reduction's carry runs serially (about a third of the
// product-scanning chain).  Same column sums, same m digits, so the result
// equals the product-scanning one bit for bit; a few more instructions.
__device__ __forceinline__ void cols29_mad(uint64_t (&t)[17], const Fq29 &a, const Fq29 &b) {
#pragma unroll
  for (int i = 0; i < 9; i++) {
#pragma unroll
    for (int j = 0; j < 9; j++) mad29(t[i + j], a.l[i], b.l[j]);
  }
}
__device__ __forceinline__ Fq29 cols29_redc(uint64_t (&t)[17]) {
  Fq29 r;
  uint64_t carry = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    t[i] += carry;
    const uint32_t m = ((uint32_t)t[i] * Q29::np) & Q29::M;
    mad29(t[i], m, Q29::p[0]);
    carry = t[i] >> 29;
#pragma unroll
    for (int j = 1; j < 9; j++) mad29(t[i + j], m, Q29::p[j]);
  }
#pragma unroll
  for (int k = 9; k < 17; k++) {
    t[k] += carry;
    r.l[k - 9] = (uint32_t)t[k] & Q29::M;
    carry = t[k] >> 29;
  }
  r.l[8] = (uint32_t)carry;
  return r;
}
__device__ __forceinline__ Fq29 mul29_lat(const Fq29 &a, const Fq29 &b) {
  uint64_t t[17] = {};
  cols29_mad(t, a, b);
  return cols29_redc(t);
}
__device__ __forceinline__ Fq29 dot29_lat(const Fq29 &a, const Fq29 &b, const Fq29 &c, const Fq29 &d) {
  uint64_t t[17] = {};
  cols29_mad(t, a, b);
  cols29_mad(t, c, d);
  return cols29_redc(t);
}

// a + K*q - b, normalised.  b's value at most K*q, its low limbs normalised.
template <int K> __device__ __forceinline__ Fq29 sub29(const Fq29 &a, const Fq29 &b) {
  Fq29 r;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const uint32_t t = a.l[i] + (QMul<K>::d[i] - b.l[i]) + c;
    if (i < 8) {
      c = t >> 29;
      r.l[i] = t & Q29::M;
    } else {
      r.l[i] = t;
    }
  }
  return r;
}

// a + k*b, normalised (k = 1 or 2)
template <int k> __device__ __forceinline__ Fq29 add29(const Fq29 &a, const Fq29 &b) {
  Fq29 r;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const uint32_t t = a.l[i] + k * b.l[i] + c;
    if (i < 8) {
      c = t >> 29;
      r.l[i] = t & Q29::M;
    } else {
      r.l[i] = t;
    }
  }
  return r;
}

// zero mod q for a normalised value below 2q: the value is 0 or q
__device__ __forceinline__ bool is_zero29(const Fq29 &a) {
  uint32_t z = a.l[0], e = a.l[0] ^ Q29::p[0];
#pragma unroll
  for (int i = 1; i < 9; i++) {
    z |= a.l[i];
    e |= a.l[i] ^ Q29::p[i];
  }
  return z == 0 || e == 0;
}

// a - m when a >= m (normalised a, m a normalised constant)
__device__ __forceinline__ Fq29 csub29(const Fq29 &a, const uint32_t (&m)[9]) {
  Fq29 d;
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int32_t t = (int32_t)a.l[i] - (int32_t)m[i] + c;
    if (i < 8) {
      c = t >> 29;  // arithmetic: 0 or -1
      d.l[i] = (uint32_t)t & Q29::M;
    } else {
      c = t;
      d.l[i] = (uint32_t)t;
    }
  }
  Fq29 r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.l[i] = c < 0 ? a.l[i] : d.l[i];
  return r;
}

// canonical representative of a normalised value below 2q
__device__ __forceinline__ Fq29 canon29(const Fq29 &a) { return csub29(a, Q29::p); }

// 8 x u32 words (a value below 2^256) -> 9 normalised limbs
__device__ __forceinline__ Fq29 unpack29(const uint32_t (&w)[8]) {
  Fq29 r;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const int b = 29 * i, wi = b >> 5, sh = b & 31;
    uint32_t x = w[wi] >> sh;
    if (sh > 3 && wi < 7) x |= w[wi + 1] << (32 - sh);
    r.l[i] = i < 8 ? (x & Q29::M) : x;
  }
  return r;
}

// 9 normalised limbs of a value below 2^256 -> 8 x u32 words
__device__ __forceinline__ void pack29(const Fq29 &a, uint32_t (&w)[8]) {
#pragma unroll
  for (int k = 0; k < 8; k++) {
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int lo = 29 * i - 32 * k;  // bit 0 of limb i relative to bit 0 of word k
      if (lo >= 32 || lo + 29 <= 0) continue;
      x |= lo >= 0 ? (a.l[i] << lo) : (a.l[i] >> -lo);
    }
    w[k] = x;
  }
}

// R = 2^256 Montgomery words (< q) -> 2^261 form, canonical, packed
__device__ __forceinline__ void to29_words(const uint32_t (&in)[8], uint32_t (&out)[8]) {
  pack29(canon29(mul29(unpack29(in), f29_const(Q29::to))), out);
}

// 2^261 form (value below 8q) -> canonical R = 2^256 words
__device__ __forceinline__ void from29_words(const Fq29 &a, uint32_t (&out)[8]) {
  pack29(canon29(mul29(a, f29_const(Q29::from))), out);
}

struct XYZZ29 {
  Fq29 X, Y, ZZ, ZZZ;
};

// 2*q for affine q (x < q, y <= q), mdbl-2008-s-1 with ranges as madd29's
// output: X3 < 6q, Y3, ZZ, ZZZ < 2q.
__device__ __forceinline__ XYZZ29 dbl29_affine(const Fq29 &x, const Fq29 &y) {
  const Fq29 U = add29<1>(y, y);
  Fq29 V, xx, W, S;
  mul29x2(U, U, x, x, V, xx);
  mul29x2(U, V, x, V, W, S);
  const Fq29 Mv = add29<2>(xx, xx);  // 3 xx < 6q
  const Fq29 X3 = sub29<4>(mul29(Mv, Mv), add29<1>(S, S));
  const Fq29 Y3 = dot29(Mv, sub29<6>(S, X3), W, sub29<2>(Fq29{}, y));  // M (S - X3) - W y
  return {X3, Y3, V, W};
}

// acc += q for affine q (x < q, y <= q), madd-2008-s.  acc ranges on entry
// and exit: X < 8q, Y, ZZ, ZZZ < 2q (bias multiples chosen from them).
// The caller tracks infinity in ainf (true exactly when acc is infinity,
// which then also has ZZ == 0): a run starts at infinity and returns there
// only by cancellation.  Updated in place, one coordinate at a time, so the
// compiler keeps acc in registers on every path.
__device__ __forceinline__ void madd29(XYZZ29 &acc, bool &ainf, const Fq29 &x, const Fq29 &y) {
  if (ainf) {
    acc.X = x;
    acc.Y = y;
    acc.ZZ = f29_const(Q29::one);
    acc.ZZZ = f29_const(Q29::one);
    ainf = false;
    return;
  }
  Fq29 U2, S2, PP, RR, PPP, Qv, ZZ3, ZZZ3;
  mul29x2(x, acc.ZZ, y, acc.ZZZ, U2, S2);
  const Fq29 P = sub29<8>(U2, acc.X);  // < 10q
  const Fq29 R = sub29<2>(S2, acc.Y);  // < 4q
  sqr29x2(P, R, PP, RR);
  if (is_zero29(PP)) {
    if (is_zero29(RR)) {
      const XYZZ29 d = dbl29_affine(x, y);
      acc.X = d.X;
      acc.Y = d.Y;
      acc.ZZ = d.ZZ;
      acc.ZZZ = d.ZZZ;
    } else {
      acc.ZZ = Fq29{};  // P + (-P): infinity
      acc.ZZZ = Fq29{};
      ainf = true;
    }
    return;
  }
  mul29x2(P, PP, acc.X, PP, PPP, Qv);
  mul29x2(acc.ZZ, PP, acc.ZZZ, PPP, ZZ3, ZZZ3);
  acc.ZZ = ZZ3;
  acc.ZZZ = ZZZ3;
  const Fq29 X3 = sub29<6>(RR, add29<2>(PPP, Qv));  // PPP + 2Qv < 6q; X3 < 8q
  // Y3 = R (Qv - X3) - Y PPP, one reduction of R T + Y (2q - PPP) < 44 q^2
  acc.Y = dot29(R, sub29<8>(Qv, X3), acc.Y, sub29<2>(Fq29{}, PPP));
  acc.X = X3;
}

// ------------------------------------------------------------------ G2
// Fq2 = Fq[u]/(u^2 + 1) on 29-bit limbs.  Each component of a product is one
// Montgomery reduction of a sum of limb products: c0 = a0 b0 + a1 (2q - b1),
// c1 = a0 b1 + a1 b0, the negated b1 supplied by the caller (shared between
// products with the same b).  The G2 loop keeps X < 4q and Y, ZZ, ZZZ < 2q
// per component; with 2^261 > 170q the squares of P = U2 - X < 6q stay in
// range (a limb-exact model of madd2_29 checks every column and range).
struct Fq2_29 {
  Fq29 c0, c1;
};
struct XYZZ2_29 {
  Fq2_29 X, Y, ZZ, ZZZ;
};

// redc(a*b + c*d) and redc(e*f + g*h): four interleaved column chains, the
// a*b and e*f products merged into their columns once per column
__device__ __forceinline__ void dot29x2(const Fq29 &a, const Fq29 &b, const Fq29 &c, const Fq29 &d,
                                        const Fq29 &e, const Fq29 &f, const Fq29 &g, const Fq29 &h, Fq29 &r1,
                                        Fq29 &r2) {
  uint32_t m[9], n[9];
  uint64_t acc = 0, acd = 0;
#pragma unroll
  for (int k = 0; k < 17; k++) {
    uint64_t u = 0, v = 0;
#pragma unroll
    for (int j = (k > 8 ? k - 8 : 0); j <= (k < 8 ? k : 8); j++) {
      mad29(u, a.l[j], b.l[k - j]);
      mad29(v, e.l[j], f.l[k - j]);
      mad29(acc, c.l[j], d.l[k - j]);
      mad29(acd, g.l[j], h.l[k - j]);
    }
#pragma unroll
    for (int j = (k > 8 ? k - 8 : 0); j < (k < 9 ? k : 9); j++) {
      mad29(acc, m[j], Q29::p[k - j]);
      mad29(acd, n[j], Q29::p[k - j]);
    }
    acc += u;
    acd += v;
    if (k < 9) {
      m[k] = ((uint32_t)acc * Q29::np) & Q29::M;
      n[k] = ((uint32_t)acd * Q29::np) & Q29::M;
      mad29(acc, m[k], Q29::p[0]);
      mad29(acd, n[k], Q29::p[0]);
    } else {
      r1.l[k - 9] = (uint32_t)acc & Q29::M;
      r2.l[k - 9] = (uint32_t)acd & Q29::M;
    }
    acc >>= 29;
    acd >>= 29;
  }
  r1.l[8] = (uint32_t)acc;
  r2.l[8] = (uint32_t)acd;
}

// redc(a0 b0 + a1 b1 + a2 b2 + a3 b3) and redc(e0 f0 + ... + e3 f3), all
// operands normalised: a column holds at most 45 products below 2^58.
__device__ __forceinline__ void dot4x2(const Fq29 &a0, const Fq29 &b0, const Fq29 &a1, const Fq29 &b1,
                                       const Fq29 &a2, const Fq29 &b2, const Fq29 &a3, const Fq29 &b3,
                                       const Fq29 &e0, const Fq29 &f0, const Fq29 &e1, const Fq29 &f1,
                                       const Fq29 &e2, const Fq29 &f2, const Fq29 &e3, const Fq29 &f3, Fq29 &r1,
                                       Fq29 &r2) {
  uint32_t m[9], n[9];
  uint64_t acc = 0, acd = 0;
#pragma unroll
  for (int k = 0; k < 17; k++) {
    uint64_t u = 0, v = 0;
#pragma unroll
    for (int j = (k > 8 ? k - 8 : 0); j <= (k < 8 ? k : 8); j++) {
      mad29(u, a0.l[j], b0.l[k - j]);
      mad29(v, e0.l[j], f0.l[k - j]);
      mad29(acc, a2.l[j], b2.l[k - j]);
      mad29(acd, e2.l[j], f2.l[k - j]);
      mad29(u, a1.l[j], b1.l[k - j]);
      mad29(v, e1.l[j], f1.l[k - j]);
      mad29(acc, a3.l[j], b3.l[k - j]);
      mad29(acd, e3.l[j], f3.l[k - j]);
    }
#pragma unroll
    for (int j = (k > 8 ? k - 8 : 0); j < (k < 9 ? k : 9); j++) {
      mad29(acc, m[j], Q29::p[k - j]);
      mad29(acd, n[j], Q29::p[k - j]);
    }
    acc += u;
    acd += v;
    if (k < 9) {
      m[k] = ((uint32_t)acc * Q29::np) & Q29::M;
      n[k] = ((uint32_t)acd * Q29::np) & Q29::M;
      mad29(acc, m[k], Q29::p[0]);
      mad29(acd, n[k], Q29::p[0]);
    } else {
      r1.l[k - 9] = (uint32_t)acc & Q29::M;
      r2.l[k - 9] = (uint32_t)acd & Q29::M;
    }
    acc >>= 29;
    acd >>= 29;
  }
  r1.l[8] = (uint32_t)acc;
  r2.l[8] = (uint32_t)acd;
}

// a*b; nb1 = Kq - b.c1 for b's component range Kq
__device__ __forceinline__ Fq2_29 f2mul29(const Fq2_29 &a, const Fq2_29 &b, const Fq29 &nb1) {
  Fq2_29 r;
  dot29x2(a.c0, b.c0, a.c1, nb1, a.c0, b.c1, a.c1, b.c0, r.c0, r.c1);
  return r;
}
// a^2 = ((a0 + a1)(a0 - a1), 2 a0 a1) for components below Kq
template <int K> __device__ __forceinline__ Fq2_29 f2sqr29(const Fq2_29 &a) {
  Fq29 s, t;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    s.l[i] = a.c0.l[i] + a.c1.l[i];  // limbs < 2^30
    t.l[i] = a.c0.l[i] << 1;
  }
  Fq2_29 r;
  mul29x2(s, sub29<K>(a.c0, a.c1), t, a.c1, r.c0, r.c1);
  return r;
}
template <int K> __device__ __forceinline__ Fq2_29 f2sub29(const Fq2_29 &a, const Fq2_29 &b) {
  return {sub29<K>(a.c0, b.c0), sub29<K>(a.c1, b.c1)};
}
__device__ __forceinline__ bool is_zero2_29(const Fq2_29 &a) { return is_zero29(a.c0) && is_zero29(a.c1); }
__device__ __forceinline__ Fq2_29 f2one29() { return {f29_const(Q29::one), Fq29{}}; }

// 2*q for affine q (components x < q, y <= q), mdbl-2008-s-1; output ranges
// as madd2_29's
__device__ __forceinline__ XYZZ2_29 dbl2_29_affine(const Fq2_29 &x, const Fq2_29 &y) {
  const Fq2_29 U = {add29<1>(y.c0, y.c0), add29<1>(y.c1, y.c1)};
  const Fq2_29 V = f2sqr29<2>(U);
  const Fq29 nV1 = sub29<2>(Fq29{}, V.c1);
  const Fq2_29 W = f2mul29(U, V, nV1);
  const Fq2_29 S = f2mul29(x, V, nV1);
  const Fq2_29 xx = f2sqr29<1>(x);
  const Fq2_29 Mv = {add29<2>(xx.c0, xx.c0), add29<2>(xx.c1, xx.c1)};  // 3 xx < 6q
  const Fq2_29 MM = f2sqr29<6>(Mv);
  Fq2_29 X3;
  X3.c0 = csub29(sub29<4>(MM.c0, add29<1>(S.c0, S.c0)), Q29::p4);
  X3.c1 = csub29(sub29<4>(MM.c1, add29<1>(S.c1, S.c1)), Q29::p4);
  const Fq2_29 T = f2sub29<4>(S, X3);  // < 6q
  const Fq29 nT1 = sub29<6>(Fq29{}, T.c1), ny0 = sub29<1>(Fq29{}, y.c0), ny1 = sub29<1>(Fq29{}, y.c1);
  Fq2_29 Y3;  // M (S - X3) - W y
  dot4x2(Mv.c0, T.c0, Mv.c1, nT1, W.c0, ny0, W.c1, y.c1, Mv.c0, T.c1, Mv.c1, T.c0, W.c0, ny1, W.c1, ny0, Y3.c0,
         Y3.c1);
  return {X3, Y3, V, W};
}

// acc += q for affine q (components x < q, y <= q), madd-2008-s over Fq2,
// with infinity tracked in ainf as for madd29.  Returns false, acc
// unchanged, when q equals acc (the caller doubles q).
__device__ __forceinline__ bool madd2_29(XYZZ2_29 &acc, bool &ainf, const Fq2_29 &x, const Fq2_29 &y) {
  if (ainf) {
    acc.X = x;
    acc.Y = y;
    acc.ZZ = f2one29();
    acc.ZZZ = f2one29();
    ainf = false;
    return true;
  }
  const Fq2_29 U2 = f2mul29(x, acc.ZZ, sub29<2>(Fq29{}, acc.ZZ.c1));
  const Fq2_29 S2 = f2mul29(y, acc.ZZZ, sub29<2>(Fq29{}, acc.ZZZ.c1));
  const Fq2_29 P = f2sub29<4>(U2, acc.X);  // < 6q
  const Fq2_29 R = f2sub29<2>(S2, acc.Y);  // < 4q
  const Fq2_29 PP = f2sqr29<6>(P);
  const Fq2_29 RR = f2sqr29<4>(R);
  if (is_zero2_29(PP)) {
    if (is_zero2_29(RR)) return false;
    acc.ZZ = Fq2_29{};  // P + (-P): infinity
    acc.ZZZ = Fq2_29{};
    ainf = true;
    return true;
  }
  const Fq29 nPP1 = sub29<2>(Fq29{}, PP.c1);
  const Fq2_29 PPP = f2mul29(P, PP, nPP1);
  const Fq2_29 Qv = f2mul29(acc.X, PP, nPP1);
  const Fq2_29 ZZ3 = f2mul29(acc.ZZ, PP, nPP1);
  const Fq29 nPPP1 = sub29<2>(Fq29{}, PPP.c1);
  const Fq2_29 ZZZ3 = f2mul29(acc.ZZZ, PPP, nPPP1);
  Fq2_29 X3;  // R^2 - PPP - 2 Qv, then below 4q
  X3.c0 = csub29(sub29<6>(RR.c0, add29<2>(PPP.c0, Qv.c0)), Q29::p4);
  X3.c1 = csub29(sub29<6>(RR.c1, add29<2>(PPP.c1, Qv.c1)), Q29::p4);
  const Fq2_29 T = f2sub29<4>(Qv, X3);  // < 6q
  const Fq29 nT1 = sub29<6>(Fq29{}, T.c1), nP0 = sub29<2>(Fq29{}, PPP.c0);
  Fq2_29 Y3;  // R T - Y PPP
  dot4x2(R.c0, T.c0, R.c1, nT1, acc.Y.c0, nP0, acc.Y.c1, PPP.c1, R.c0, T.c1, R.c1, T.c0, acc.Y.c0, nPPP1, acc.Y.c1,
         nP0, Y3.c0, Y3.c1);
  acc.X = X3;
  acc.Y = Y3;
  acc.ZZ = ZZ3;
  acc.ZZZ = ZZZ3;
  return true;
}

// ------------------------------------------------------------------ general
// addition and doubling on stored 29-bit slots (carries, bucket and window
// reductions).  Inputs and outputs keep the loops' ranges: G1 X < 8q, G2
// X < 4q per component, Y, ZZ, ZZZ < 2q; infinity is ZZ == 0 mod q.

// a = 2a (dbl-2008-s, a = 0)
__device__ __forceinline__ void xdbl29(XYZZ29 &a) {
  if (is_zero29(a.ZZ)) return;
  const Fq29 U = add29<1>(a.Y, a.Y);  // < 4q
  Fq29 V, xx, W, S, MM, ZZ3;
  sqr29x2(U, a.X, V, xx);
  mul29x2(U, V, a.X, V, W, S);
  const Fq29 Mv = add29<2>(xx, xx);  // 3 xx < 6q
  mul29x2(Mv, Mv, V, a.ZZ, MM, ZZ3);
  const Fq29 X3 = sub29<4>(MM, add29<1>(S, S));                           // < 6q
  const Fq29 Y3 = dot29(Mv, sub29<6>(S, X3), W, sub29<2>(Fq29{}, a.Y));  // M (S - X3) - W Y
  a.ZZZ = mul29(W, a.ZZZ);
  a.ZZ = ZZ3;
  a.X = X3;
  a.Y = Y3;
}

// a += b (add-2008-s)
__device__ __forceinline__ void xadd29(XYZZ29 &a, const XYZZ29 &b) {
  if (is_zero29(b.ZZ)) return;
  if (is_zero29(a.ZZ)) {
    a = b;
    return;
  }
  Fq29 U1, U2, S1, S2, PP, RR, PPP, Qv, Z2, Z3, ZZ3, ZZZ3;
  mul29x2(a.X, b.ZZ, b.X, a.ZZ, U1, U2);
  mul29x2(a.Y, b.ZZZ, b.Y, a.ZZZ, S1, S2);
  const Fq29 P = sub29<2>(U2, U1), R = sub29<2>(S2, S1);  // < 4q
  sqr29x2(P, R, PP, RR);
  if (is_zero29(PP)) {
    if (is_zero29(RR)) {
      xdbl29(a);
    } else {
      a.ZZ = Fq29{};
      a.ZZZ = Fq29{};
    }
    return;
  }
  mul29x2(P, PP, U1, PP, PPP, Qv);
  mul29x2(a.ZZ, b.ZZ, a.ZZZ, b.ZZZ, Z2, Z3);
  mul29x2(Z2, PP, Z3, PPP, ZZ3, ZZZ3);
  const Fq29 X3 = sub29<6>(RR, add29<2>(PPP, Qv));                          // < 8q
  a.Y = dot29(R, sub29<8>(Qv, X3), S1, sub29<2>(Fq29{}, PPP));            // R (Qv - X3) - S1 PPP
  a.X = X3;
  a.ZZ = ZZ3;
  a.ZZZ = ZZZ3;
}

// b with its second component negated (2q - b.c1), for f2mul29
__device__ __forceinline__ Fq2_29 f2mul29n(const Fq2_29 &a, const Fq2_29 &b) {
  return f2mul29(a, b, sub29<2>(Fq29{}, b.c1));
}

__device__ __forceinline__ void xdbl29(XYZZ2_29 &a) {
  if (is_zero2_29(a.ZZ)) return;
  const Fq2_29 U = {add29<1>(a.Y.c0, a.Y.c0), add29<1>(a.Y.c1, a.Y.c1)};  // < 4q
  const Fq2_29 V = f2sqr29<4>(U);
  const Fq29 nV1 = sub29<2>(Fq29{}, V.c1);
  const Fq2_29 W = f2mul29(U, V, nV1);
  const Fq2_29 S = f2mul29(a.X, V, nV1);
  const Fq2_29 xx = f2sqr29<4>(a.X);
  const Fq2_29 Mv = {add29<2>(xx.c0, xx.c0), add29<2>(xx.c1, xx.c1)};  // 3 xx < 6q
  const Fq2_29 MM = f2sqr29<6>(Mv);
  Fq2_29 X3;
  X3.c0 = csub29(sub29<4>(MM.c0, add29<1>(S.c0, S.c0)), Q29::p4);
  X3.c1 = csub29(sub29<4>(MM.c1, add29<1>(S.c1, S.c1)), Q29::p4);
  const Fq2_29 T = f2sub29<4>(S, X3);  // < 6q
  const Fq29 nT1 = sub29<6>(Fq29{}, T.c1), ny0 = sub29<2>(Fq29{}, a.Y.c0), ny1 = sub29<2>(Fq29{}, a.Y.c1);
  Fq2_29 Y3;  // M (S - X3) - W Y
  dot4x2(Mv.c0, T.c0, Mv.c1, nT1, W.c0, ny0, W.c1, a.Y.c1, Mv.c0, T.c1, Mv.c1, T.c0, W.c0, ny1, W.c1, ny0, Y3.c0,
         Y3.c1);
  a.ZZ = f2mul29n(V, a.ZZ);
  a.ZZZ = f2mul29n(W, a.ZZZ);
  a.X = X3;
  a.Y = Y3;
}

__device__ __forceinline__ void xadd29(XYZZ2_29 &a, const XYZZ2_29 &b) {
  if (is_zero2_29(b.ZZ)) return;
  if (is_zero2_29(a.ZZ)) {
    a = b;
    return;
  }
  const Fq29 nbZZ1 = sub29<2>(Fq29{}, b.ZZ.c1), nbZZZ1 = sub29<2>(Fq29{}, b.ZZZ.c1);
  const Fq2_29 U1 = f2mul29(a.X, b.ZZ, nbZZ1);
  const Fq2_29 U2 = f2mul29n(b.X, a.ZZ);
  const Fq2_29 S1 = f2mul29(a.Y, b.ZZZ, nbZZZ1);
  const Fq2_29 S2 = f2mul29n(b.Y, a.ZZZ);
  const Fq2_29 P = f2sub29<2>(U2, U1), R = f2sub29<2>(S2, S1);  // < 4q
  const Fq2_29 PP = f2sqr29<4>(P);
  const Fq2_29 RR = f2sqr29<4>(R);
  if (is_zero2_29(PP)) {
    if (is_zero2_29(RR)) {
      xdbl29(a);
    } else {
      a.ZZ = Fq2_29{};
      a.ZZZ = Fq2_29{};
    }
    return;
  }
  const Fq29 nPP1 = sub29<2>(Fq29{}, PP.c1);
  const Fq2_29 PPP = f2mul29(P, PP, nPP1);
  const Fq2_29 Qv = f2mul29(U1, PP, nPP1);
  const Fq29 nPPP1 = sub29<2>(Fq29{}, PPP.c1);
  const Fq2_29 ZZ3 = f2mul29(f2mul29(a.ZZ, b.ZZ, nbZZ1), PP, nPP1);
  const Fq2_29 ZZZ3 = f2mul29(f2mul29(a.ZZZ, b.ZZZ, nbZZZ1), PPP, nPPP1);
  Fq2_29 X3;
  X3.c0 = csub29(sub29<6>(RR.c0, add29<2>(PPP.c0, Qv.c0)), Q29::p4);
  X3.c1 = csub29(sub29<6>(RR.c1, add29<2>(PPP.c1, Qv.c1)), Q29::p4);
  const Fq2_29 T = f2sub29<4>(Qv, X3);  // < 6q
  const Fq29 nT1 = sub29<6>(Fq29{}, T.c1), nP0 = sub29<2>(Fq29{}, PPP.c0);
  Fq2_29 Y3;  // R T - S1 PPP
  dot4x2(R.c0, T.c0, R.c1, nT1, S1.c0, nP0, S1.c1, PPP.c1, R.c0, T.c1, R.c1, T.c0, S1.c0, nPPP1, S1.c1, nP0, Y3.c0,
         Y3.c1);
  a.X = X3;
  a.Y = Y3;
  a.ZZ = ZZ3;
  a.ZZZ = ZZZ3;
}

// k a for a small k (bucket-segment offsets), left to right
template <class A> __device__ __forceinline__ A xmul_small29(const A &p, uint32_t k, const A &inf) {
  if (k == 0) return inf;
  A acc = p;
  const int top = 31 - __builtin_clz(k);
#pragma unroll 1
  for (int b = top - 1; b >= 0; b--) {
    xdbl29(acc);
    if ((k >> b) & 1u) xadd29(acc, p);
  }
  return acc;
}

}  // namespace csk
