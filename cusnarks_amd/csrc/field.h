// BN254 base/scalar field arithmetic for gfx950 (CDNA4).
//
// Elements are 8 x u32 little-endian limbs in Montgomery form (R = 2^256),
// the layout the reference uses everywhere (cusnarks src/cuda/types.h:150,
// constants.cpp:49-67).  PIDX 0 is the base field q (coordinates), PIDX 1 the
// scalar field r, matching the reference's MOD_GROUP / MOD_FIELD indices.
//
// Multiplication is Montgomery CIOS over 32-bit limbs, written so that every
// limb product lands on v_mad_u64_u32 (a 32x32+64 -> 64 multiply-add) with
// the running carry folded into the 64-bit addend.  Both moduli have a top
// limb below 2^31 - 1, so the "no final carry word" CIOS variant applies: the
// accumulator never needs a ninth limb and every product result is < 2p,
// brought to [0, p) by one conditional subtraction.  All results are fully
// reduced, so equal field elements are equal bit patterns (the reference's
// montmult_h, cusnarks src/cuda/ff.cpp:225-388, is fully reduced as well).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace csk {

template <int PIDX> struct FieldConst;

template <> struct FieldConst<0> {  // q
  static constexpr uint32_t p[8] = {3632069959u, 1008765974u, 1752287885u, 2541841041u,
                                    2172737629u, 3092268470u, 3778125865u, 811880050u};
  static constexpr uint32_t np = 3834012553u;  // -q^-1 mod 2^32
  static constexpr uint32_t one[8] = {3314486685u, 3546104717u, 4123462461u, 175696680u,
                                      2021213740u, 1718526831u, 2584207151u, 235567041u};
  static constexpr uint32_t r2[8] = {1401617033u, 4079811675u, 3561292283u, 3051821329u,
                                     172064758u,  1202396927u, 3401069855u, 114859889u};
  static constexpr uint32_t r3[8] = {3658821855u, 2983030191u, 2804432310u, 1660031206u,
                                     182061819u,  4018080524u, 760816964u,  553479824u};
};

template <> struct FieldConst<1> {  // r
  static constexpr uint32_t p[8] = {4026531841u, 1138881939u, 2042196113u, 674490440u,
                                    2172737629u, 3092268470u, 3778125865u, 811880050u};
  static constexpr uint32_t np = 4026531839u;
  static constexpr uint32_t one[8] = {1342177275u, 2895524892u, 2673921321u, 922515093u,
                                      2021213742u, 1718526831u, 2584207151u, 235567041u};
  static constexpr uint32_t r2[8] = {2921426343u, 465102405u,  3814480355u, 1409170097u,
                                     1404797061u, 2353627965u, 2135835813u, 35049649u};
  static constexpr uint32_t r3[8] = {3032416320u, 1586813153u, 486258360u,  709401790u,
                                     2711605229u, 2302461540u, 2144101756u, 217602379u};
};

template <int PIDX> struct Fp {
  uint32_t v[8];

  using C = FieldConst<PIDX>;

  __device__ __forceinline__ static Fp zero() {
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = 0;
    return r;
  }
  __device__ __forceinline__ static Fp one() {
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = C::one[i];
    return r;
  }
  __device__ __forceinline__ bool is_zero() const {
    uint32_t acc = v[0];
#pragma unroll
    for (int i = 1; i < 8; i++) acc |= v[i];
    return acc == 0;
  }
  __device__ __forceinline__ bool is_one() const {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) acc |= v[i] ^ C::one[i];
    return acc == 0;
  }
  __device__ __forceinline__ bool operator==(const Fp &o) const {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) acc |= v[i] ^ o.v[i];
    return acc == 0;
  }

  // x - p if x >= p (x < 2p on entry)
  __device__ __forceinline__ static Fp reduce_once(const Fp &x) {
    // limb chains on the carry flag (v_sub_co / v_subb_co): one instruction
    // per limb, where 64-bit temporaries cost three to five
    Fp d;
    unsigned borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] = __builtin_subc(x.v[i], C::p[i], borrow, &borrow);
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = borrow ? x.v[i] : d.v[i];
    return r;
  }

  __device__ __forceinline__ static Fp add(const Fp &a, const Fp &b) {
    // p < 2^254, so a + b < 2^255 fits in 8 limbs
    Fp s;
    unsigned carry = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s.v[i] = __builtin_addc(a.v[i], b.v[i], carry, &carry);
    return reduce_once(s);
  }

  __device__ __forceinline__ static Fp sub(const Fp &a, const Fp &b) {
    Fp d;
    unsigned borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] = __builtin_subc(a.v[i], b.v[i], borrow, &borrow);
    // add p back when the subtraction wrapped
    const uint32_t mask = 0u - (uint32_t)borrow;
    unsigned carry = 0;
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = __builtin_addc(d.v[i], C::p[i] & mask, carry, &carry);
    return r;
  }

  // Sums on the lazy range [0, 2p) (the transforms' butterflies, whose
  // products are mul_lazy): 4p < 2^256, so a + b needs one conditional
  // subtraction of 2p and a - b one conditional addition of 2p.
  struct Limbs {
    uint32_t v[8];
  };
  static constexpr Limbs twice_p() {
    Limbs r = {};
    uint64_t c = 0;
    for (int i = 0; i < 8; i++) {
      const uint64_t t = 2ull * C::p[i] + c;
      r.v[i] = (uint32_t)t;
      c = t >> 32;
    }
    return r;
  }
  __device__ __forceinline__ static Fp add_lazy(const Fp &a, const Fp &b) {
    constexpr Limbs p2 = twice_p();
    Fp s, d;
    unsigned carry = 0, borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s.v[i] = __builtin_addc(a.v[i], b.v[i], carry, &carry);
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] = __builtin_subc(s.v[i], p2.v[i], borrow, &borrow);
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = borrow ? s.v[i] : d.v[i];
    return r;
  }
  __device__ __forceinline__ static Fp sub_lazy(const Fp &a, const Fp &b) {
    constexpr Limbs p2 = twice_p();
    Fp d;
    unsigned borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] = __builtin_subc(a.v[i], b.v[i], borrow, &borrow);
    const uint32_t mask = 0u - (uint32_t)borrow;
    unsigned carry = 0;
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = __builtin_addc(d.v[i], p2.v[i] & mask, carry, &carry);
    return r;
  }

  __device__ __forceinline__ static Fp neg(const Fp &a) { return sub(zero(), a); }
  __device__ __forceinline__ static Fp dbl(const Fp &a) { return add(a, a); }

  // One multiply-accumulate step of the product-scanning product:
  // {c2:acc} += x*y, with v_mad_u64_u32 (32x32 + 64 -> 64, carry-out in VCC)
  // and the carry folded into the third column word by v_addc_co_u32.
  __device__ __forceinline__ static void mac(uint64_t &acc, uint32_t &c2, uint32_t x, uint32_t y) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(c2)
        : "v"(x), "v"(y)
        : "vcc");
  }
  // first product of a column: c2 = carry-out (no prior c2 to add to)
  __device__ __forceinline__ static void mac_first(uint64_t &acc, uint32_t &c2, uint32_t x, uint32_t y) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_cndmask_b32_e64 %1, 0, 1, vcc"
        : "+v"(acc), "=v"(c2)
        : "v"(x), "v"(y)
        : "vcc");
  }
  // same with y in an SGPR (the uniform modulus limbs)
  __device__ __forceinline__ static void mac_s(uint64_t &acc, uint32_t &c2, uint32_t x, uint32_t y) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(c2)
        : "v"(x), "s"(y)
        : "vcc");
  }

  // Montgomery product a*b*2^-256 mod p by finely integrated product scanning:
  // column k accumulates sum_j a_j b_(k-j) + sum_j m_j p_(k-j) in a 96-bit
  // {c2:acc}; for k < 8 the reduction digit m_k = acc_lo * (-p^-1) clears the
  // low word, for k >= 8 the low word is result limb k-8.  128 limb products.
  // For a, b < 2p the result is < 2p (the lazy range); mul subtracts p once
  // more for the canonical result.
  __device__ __forceinline__ static Fp mul_lazy(const Fp &a, const Fp &b) {
    uint32_t m[8], r[8];
    uint64_t acc = 0;
    uint32_t c2 = 0;
#pragma unroll
    for (int k = 0; k < 15; k++) {
      const int j0 = k > 7 ? k - 7 : 0;
      mac_first(acc, c2, a.v[j0], b.v[k - j0]);
#pragma unroll
      for (int j = j0 + 1; j <= (k < 7 ? k : 7); j++) mac(acc, c2, a.v[j], b.v[k - j]);
#pragma unroll
      for (int j = (k > 7 ? k - 7 : 0); j < (k < 8 ? k : 8); j++) mac_s(acc, c2, m[j], C::p[k - j]);
      if (k < 8) {
        m[k] = (uint32_t)acc * C::np;
        mac_s(acc, c2, m[k], C::p[0]);
      } else {
        r[k - 8] = (uint32_t)acc;
      }
      acc = (acc >> 32) | ((uint64_t)c2 << 32);
    }
    r[7] = (uint32_t)acc;
    Fp out;
#pragma unroll
    for (int i = 0; i < 8; i++) out.v[i] = r[i];
    return out;
  }
  __device__ __forceinline__ static Fp mul(const Fp &a, const Fp &b) { return reduce_once(mul_lazy(a, b)); }

  // Reference CIOS formulation (compiler-scheduled), kept for the VALU probe.
  __device__ __forceinline__ static Fp mul_cios(const Fp &a, const Fp &b) {
    uint32_t t[8];
#pragma unroll
    for (int i = 0; i < 8; i++) t[i] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      uint64_t A = (uint64_t)a.v[0] * b.v[i] + t[0];
      t[0] = (uint32_t)A;
      uint32_t m = t[0] * C::np;
      uint64_t Cc = (uint64_t)m * C::p[0] + t[0];
#pragma unroll
      for (int j = 1; j < 8; j++) {
        A = (uint64_t)a.v[j] * b.v[i] + ((uint64_t)t[j] + (A >> 32));
        t[j] = (uint32_t)A;
        Cc = (uint64_t)m * C::p[j] + ((uint64_t)t[j] + (Cc >> 32));
        t[j - 1] = (uint32_t)Cc;
      }
      t[7] = (uint32_t)(Cc >> 32) + (uint32_t)(A >> 32);
    }
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = t[i];
    return reduce_once(r);
  }

  __device__ __forceinline__ static Fp sqr(const Fp &a) { return mul(a, a); }

  __device__ __forceinline__ static Fp to_mont(const Fp &a) {
    Fp r2;
#pragma unroll
    for (int i = 0; i < 8; i++) r2.v[i] = C::r2[i];
    return mul(a, r2);
  }
  __device__ __forceinline__ static Fp from_mont(const Fp &a) {
    Fp one_n = zero();
    one_n.v[0] = 1;
    return mul(a, one_n);
  }

  // Inverse in the Montgomery domain: (aR)^-1 R.  Binary extended Euclid on
  // the integer aR (short latency, branchy: meant for the handful of
  // single-thread inversions that close an MSM), then a product with R^3.
  // Returns 0 for a == 0.
  __device__ static Fp inv(const Fp &a) {
    if (a.is_zero()) return zero();
    uint32_t u[8], w[8], x1[8], x2[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      u[i] = a.v[i];
      w[i] = C::p[i];
      x1[i] = 0;
      x2[i] = 0;
    }
    x1[0] = 1;
    // invariant: x1*a == u, x2*a == w (mod p); stop when u == 1 or w == 1
    for (;;) {
      bool u1 = (u[0] == 1), w1 = (w[0] == 1);
      uint32_t hu = 0, hw = 0;
      for (int i = 1; i < 8; i++) { hu |= u[i]; hw |= w[i]; }
      if (u1 && hu == 0) break;
      if (w1 && hw == 0) {
        for (int i = 0; i < 8; i++) x1[i] = x2[i];
        break;
      }
      while ((u[0] & 1) == 0) {
        shr1(u);
        half_mod(x1);
      }
      while ((w[0] & 1) == 0) {
        shr1(w);
        half_mod(x2);
      }
      if (geq(u, w)) {
        sub_raw(u, w);
        sub_mod_raw(x1, x2);
      } else {
        sub_raw(w, u);
        sub_mod_raw(x2, x1);
      }
    }
    Fp r, r3;
    for (int i = 0; i < 8; i++) {
      r.v[i] = x1[i];
      r3.v[i] = C::r3[i];
    }
    return mul(r, r3);  // (aR)^-1 * R^3 * R^-1 = a^-1 R
  }

 private:
  __device__ __forceinline__ static void shr1(uint32_t *x) {
    for (int i = 0; i < 7; i++) x[i] = (x[i] >> 1) | (x[i + 1] << 31);
    x[7] >>= 1;
  }
  // x/2 mod p for x < p
  __device__ __forceinline__ static void half_mod(uint32_t *x) {
    if (x[0] & 1) {
      uint32_t carry = 0;
      for (int i = 0; i < 8; i++) {
        uint64_t t = (uint64_t)x[i] + C::p[i] + carry;
        x[i] = (uint32_t)t;
        carry = (uint32_t)(t >> 32);
      }
      // x + p < 2^255: no ninth bit
    }
    shr1(x);
  }
  __device__ __forceinline__ static bool geq(const uint32_t *a, const uint32_t *b) {
    for (int i = 7; i >= 0; i--) {
      if (a[i] != b[i]) return a[i] > b[i];
    }
    return true;
  }
  __device__ __forceinline__ static void sub_raw(uint32_t *a, const uint32_t *b) {
    uint32_t borrow = 0;
    for (int i = 0; i < 8; i++) {
      uint64_t t = (uint64_t)a[i] - b[i] - borrow;
      a[i] = (uint32_t)t;
      borrow = (uint32_t)(t >> 63);
    }
  }
  // a = a - b mod p (a, b < p)
  __device__ __forceinline__ static void sub_mod_raw(uint32_t *a, const uint32_t *b) {
    uint32_t borrow = 0;
    for (int i = 0; i < 8; i++) {
      uint64_t t = (uint64_t)a[i] - b[i] - borrow;
      a[i] = (uint32_t)t;
      borrow = (uint32_t)(t >> 63);
    }
    if (borrow) {
      uint32_t carry = 0;
      for (int i = 0; i < 8; i++) {
        uint64_t t = (uint64_t)a[i] + C::p[i] + carry;
        a[i] = (uint32_t)t;
        carry = (uint32_t)(t >> 32);
      }
    }
  }
};

using Fq = Fp<0>;
using Fr = Fp<1>;

// Fq2 = Fq[u]/(u^2 + 1): the G2 coordinate field (nonresidue q - 1,
// cusnarks src/cuda/constants.cpp:175; mulmontu256_2 u256_device.cu:612-648).
struct Fq2 {
  Fq c0, c1;

  __device__ __forceinline__ static Fq2 zero() { return {Fq::zero(), Fq::zero()}; }
  __device__ __forceinline__ static Fq2 one() { return {Fq::one(), Fq::zero()}; }
  __device__ __forceinline__ bool is_zero() const { return c0.is_zero() && c1.is_zero(); }
  __device__ __forceinline__ bool is_one() const { return c0.is_one() && c1.is_zero(); }
  __device__ __forceinline__ bool operator==(const Fq2 &o) const { return c0 == o.c0 && c1 == o.c1; }
  __device__ __forceinline__ static Fq2 add(const Fq2 &a, const Fq2 &b) {
    return {Fq::add(a.c0, b.c0), Fq::add(a.c1, b.c1)};
  }
  __device__ __forceinline__ static Fq2 sub(const Fq2 &a, const Fq2 &b) {
    return {Fq::sub(a.c0, b.c0), Fq::sub(a.c1, b.c1)};
  }
  __device__ __forceinline__ static Fq2 neg(const Fq2 &a) { return {Fq::neg(a.c0), Fq::neg(a.c1)}; }
  __device__ __forceinline__ static Fq2 dbl(const Fq2 &a) { return add(a, a); }
  // Karatsuba: 3 base-field products
  __device__ __forceinline__ static Fq2 mul(const Fq2 &a, const Fq2 &b) {
    Fq t0 = Fq::mul(a.c0, b.c0);
    Fq t1 = Fq::mul(a.c1, b.c1);
    Fq s = Fq::mul(Fq::add(a.c0, a.c1), Fq::add(b.c0, b.c1));
    return {Fq::sub(t0, t1), Fq::sub(Fq::sub(s, t0), t1)};
  }
  // (a0 + a1 u)^2 = (a0 + a1)(a0 - a1) + 2 a0 a1 u
  __device__ __forceinline__ static Fq2 sqr(const Fq2 &a) {
    Fq m = Fq::mul(a.c0, a.c1);
    return {Fq::mul(Fq::add(a.c0, a.c1), Fq::sub(a.c0, a.c1)), Fq::dbl(m)};
  }
  __device__ static Fq2 inv(const Fq2 &a) {
    Fq d = Fq::inv(Fq::add(Fq::sqr(a.c0), Fq::sqr(a.c1)));
    return {Fq::mul(a.c0, d), Fq::neg(Fq::mul(a.c1, d))};
  }
};

}  // namespace csk
