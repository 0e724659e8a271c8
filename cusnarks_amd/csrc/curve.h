// G1 / G2 point arithmetic on y^2 = x^3 + b for gfx950.
//
// Buckets and running sums use extended-Jacobian "XYZZ" coordinates
// (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2): a mixed affine+XYZZ addition costs
// 8M+2S, a general addition 12M+2S and a doubling 6M+3S, and none needs the
// curve constant (a = 0).  Infinity is ZZ == 0.  Every formula handles the
// exceptional cases exactly (P + P, P + (-P), either input at infinity), as
// the reference formulas do (cusnarks src/cuda/ec.cpp:131-293, 462-514,
// ecbn128_device.cu:1086-1457), so results are exact group elements and the
// final affine output is canonical.
#pragma once
#include <type_traits>
#include "field.h"

namespace csk {

template <class F> struct Affine {
  F x, y;
};

template <class F> struct XYZZ {
  F X, Y, ZZ, ZZZ;
  __device__ __forceinline__ static XYZZ inf() { return {F::zero(), F::one(), F::zero(), F::zero()}; }
  __device__ __forceinline__ bool is_inf() const { return ZZ.is_zero(); }
  __device__ __forceinline__ static XYZZ from_affine(const Affine<F> &p) { return {p.x, p.y, F::one(), F::one()}; }
};

// 2*P for affine P (mdbl-2008-s-1); y != 0 always holds on BN254 (odd order).
template <class F> __device__ __forceinline__ XYZZ<F> xyzz_dbl_affine_inl(const Affine<F> &p) {
  F U = F::dbl(p.y);
  F V = F::sqr(U);
  F W = F::mul(U, V);
  F S = F::mul(p.x, V);
  F xx = F::sqr(p.x);
  F M = F::add(F::dbl(xx), xx);
  F X3 = F::sub(F::sqr(M), F::dbl(S));
  F Y3 = F::sub(F::mul(M, F::sub(S, X3)), F::mul(W, p.y));
  return {X3, Y3, V, W};
}

// 2*P (dbl-2008-s-1)
template <class F> __device__ __forceinline__ XYZZ<F> xyzz_dbl_inl(const XYZZ<F> &p) {
  if (p.is_inf()) return p;
  F U = F::dbl(p.Y);
  F V = F::sqr(U);
  F W = F::mul(U, V);
  F S = F::mul(p.X, V);
  F xx = F::sqr(p.X);
  F M = F::add(F::dbl(xx), xx);
  F X3 = F::sub(F::sqr(M), F::dbl(S));
  F Y3 = F::sub(F::mul(M, F::sub(S, X3)), F::mul(W, p.Y));
  return {X3, Y3, F::mul(V, p.ZZ), F::mul(W, p.ZZZ)};
}

// acc + q, q affine (madd-2008-s); xyzz_madd_inl is the always-inlined body,
// xyzz_madd the dispatching entry point.
template <class F> __device__ __forceinline__ XYZZ<F> xyzz_madd_inl(const XYZZ<F> &acc, const Affine<F> &q) {
  if (acc.is_inf()) return XYZZ<F>::from_affine(q);
  F U2 = F::mul(q.x, acc.ZZ);
  F S2 = F::mul(q.y, acc.ZZZ);
  F P = F::sub(U2, acc.X);
  F R = F::sub(S2, acc.Y);
  if (P.is_zero()) {
    if (R.is_zero()) return xyzz_dbl_affine(q);
    return XYZZ<F>::inf();
  }
  F PP = F::sqr(P);
  F PPP = F::mul(P, PP);
  F Qv = F::mul(acc.X, PP);
  F X3 = F::sub(F::sub(F::sqr(R), PPP), F::dbl(Qv));
  F Y3 = F::sub(F::mul(R, F::sub(Qv, X3)), F::mul(acc.Y, PPP));
  return {X3, Y3, F::mul(acc.ZZ, PP), F::mul(acc.ZZZ, PPP)};
}



// a + b (add-2008-s); xyzz_add_inl is the always-inlined body
template <class F> __device__ __forceinline__ XYZZ<F> xyzz_add_inl(const XYZZ<F> &a, const XYZZ<F> &b) {
  if (a.is_inf()) return b;
  if (b.is_inf()) return a;
  F U1 = F::mul(a.X, b.ZZ);
  F U2 = F::mul(b.X, a.ZZ);
  F S1 = F::mul(a.Y, b.ZZZ);
  F S2 = F::mul(b.Y, a.ZZZ);
  F P = F::sub(U2, U1);
  F R = F::sub(S2, S1);
  if (P.is_zero()) {
    if (R.is_zero()) return xyzz_dbl(a);
    return XYZZ<F>::inf();
  }
  F PP = F::sqr(P);
  F PPP = F::mul(P, PP);
  F Qv = F::mul(U1, PP);
  F X3 = F::sub(F::sub(F::sqr(R), PPP), F::dbl(Qv));
  F Y3 = F::sub(F::mul(R, F::sub(Qv, X3)), F::mul(S1, PPP));
  return {X3, Y3, F::mul(F::mul(a.ZZ, b.ZZ), PP), F::mul(F::mul(a.ZZZ, b.ZZZ), PPP)};
}


// k*p for a small non-negative k (bucket-segment offsets), left to right
template <class F> __device__ __forceinline__ XYZZ<F> xyzz_mul_small_inl(const XYZZ<F> &p, uint32_t k) {
  XYZZ<F> acc = XYZZ<F>::inf();
  if (k == 0 || p.is_inf()) return acc;
  int top = 31 - __builtin_clz(k);
  acc = p;
  for (int b = top - 1; b >= 0; b--) {
    acc = xyzz_dbl(acc);
    if ((k >> b) & 1u) acc = xyzz_add(acc, p);
  }
  return acc;
}

// canonical affine; infinity -> (0, 1) as the reference encodes it
// (ec_jac2aff_h, cusnarks src/cuda/ec.cpp:963-1012)
template <class F> __device__ __forceinline__ Affine<F> xyzz_to_affine_inl(const XYZZ<F> &p) {
  if (p.is_inf()) return {F::zero(), F::one()};
  F t = F::inv(F::mul(p.ZZ, p.ZZZ));  // 1/(ZZ ZZZ)
  F izz = F::mul(t, p.ZZZ);           // 1/ZZ
  F izzz = F::mul(t, p.ZZ);           // 1/ZZZ
  return {F::mul(p.X, izz), F::mul(p.Y, izzz)};
}

// Dispatching entry points: G1 (Fq) code gets the formulas inlined (no call
// frames, no scratch); G2 (Fq2) code calls one out-of-line copy per formula,
// which keeps its kernels' register count and compile time in check.
template <class F> __device__ __noinline__ XYZZ<F> xyzz_dbl_affine_call(const Affine<F> &p) { return xyzz_dbl_affine_inl(p); }
template <class F> __device__ __noinline__ XYZZ<F> xyzz_dbl_call(const XYZZ<F> &p) { return xyzz_dbl_inl(p); }
template <class F> __device__ __noinline__ XYZZ<F> xyzz_madd_call(const XYZZ<F> &a, const Affine<F> &q) { return xyzz_madd_inl(a, q); }
template <class F> __device__ __noinline__ XYZZ<F> xyzz_add_call(const XYZZ<F> &a, const XYZZ<F> &b) { return xyzz_add_inl(a, b); }
template <class F> __device__ __noinline__ XYZZ<F> xyzz_mul_small_call(const XYZZ<F> &p, uint32_t k) { return xyzz_mul_small_inl(p, k); }
template <class F> __device__ __noinline__ Affine<F> xyzz_to_affine_call(const XYZZ<F> &p) { return xyzz_to_affine_inl(p); }

#define CSK_DISPATCH(F, name, ...) \
  if constexpr (std::is_same<F, Fq>::value) return name##_inl(__VA_ARGS__); \
  else return name##_call(__VA_ARGS__)

template <class F> __device__ __forceinline__ XYZZ<F> xyzz_dbl_affine(const Affine<F> &p) { CSK_DISPATCH(F, xyzz_dbl_affine, p); }
template <class F> __device__ __forceinline__ XYZZ<F> xyzz_dbl(const XYZZ<F> &p) { CSK_DISPATCH(F, xyzz_dbl, p); }
template <class F> __device__ __forceinline__ XYZZ<F> xyzz_madd(const XYZZ<F> &a, const Affine<F> &q) { CSK_DISPATCH(F, xyzz_madd, a, q); }
template <class F> __device__ __forceinline__ XYZZ<F> xyzz_add(const XYZZ<F> &a, const XYZZ<F> &b) { CSK_DISPATCH(F, xyzz_add, a, b); }
template <class F> __device__ __forceinline__ XYZZ<F> xyzz_mul_small(const XYZZ<F> &p, uint32_t k) { CSK_DISPATCH(F, xyzz_mul_small, p, k); }
template <class F> __device__ __forceinline__ Affine<F> xyzz_to_affine(const XYZZ<F> &p) { CSK_DISPATCH(F, xyzz_to_affine, p); }

// ------------------------------------------------------------------ memory I/O
// Point layouts in HBM follow the reference: G1 affine [X | Y] (16 u32),
// G2 affine [X.c0 | X.c1 | Y.c0 | Y.c1] (32 u32).

__device__ __forceinline__ void ld8(uint32_t *d, const uint32_t *s) {
  const uint4 *s4 = reinterpret_cast<const uint4 *>(s);
  uint4 a = s4[0], b = s4[1];
  d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w;
  d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
}
__device__ __forceinline__ void st8(uint32_t *d, const uint32_t *s) {
  uint4 *d4 = reinterpret_cast<uint4 *>(d);
  d4[0] = make_uint4(s[0], s[1], s[2], s[3]);
  d4[1] = make_uint4(s[4], s[5], s[6], s[7]);
}

template <class F> struct FieldIO;
template <> struct FieldIO<Fq> {
  static constexpr int NW = 8;
  __device__ __forceinline__ static Fq load(const uint32_t *p) { Fq r; ld8(r.v, p); return r; }
  __device__ __forceinline__ static void store(uint32_t *p, const Fq &a) { st8(p, a.v); }
};
template <> struct FieldIO<Fq2> {
  static constexpr int NW = 16;
  __device__ __forceinline__ static Fq2 load(const uint32_t *p) {
    Fq2 r;
    ld8(r.c0.v, p);
    ld8(r.c1.v, p + 8);
    return r;
  }
  __device__ __forceinline__ static void store(uint32_t *p, const Fq2 &a) {
    st8(p, a.c0.v);
    st8(p + 8, a.c1.v);
  }
};

template <class F> __device__ __forceinline__ Affine<F> load_affine(const uint32_t *p) {
  constexpr int NW = FieldIO<F>::NW;
  return {FieldIO<F>::load(p), FieldIO<F>::load(p + NW)};
}
template <class F> __device__ __forceinline__ void store_affine(uint32_t *p, const Affine<F> &a) {
  constexpr int NW = FieldIO<F>::NW;
  FieldIO<F>::store(p, a.x);
  FieldIO<F>::store(p + NW, a.y);
}
template <class F> __device__ __forceinline__ XYZZ<F> load_xyzz(const uint32_t *p) {
  constexpr int NW = FieldIO<F>::NW;
  return {FieldIO<F>::load(p), FieldIO<F>::load(p + NW), FieldIO<F>::load(p + 2 * NW),
          FieldIO<F>::load(p + 3 * NW)};
}
template <class F> __device__ __forceinline__ void store_xyzz(uint32_t *p, const XYZZ<F> &a) {
  constexpr int NW = FieldIO<F>::NW;
  FieldIO<F>::store(p, a.X);
  FieldIO<F>::store(p + NW, a.Y);
  FieldIO<F>::store(p + 2 * NW, a.ZZ);
  FieldIO<F>::store(p + 3 * NW, a.ZZZ);
}

// Affine input at infinity: X == 0 and Y in {0, 1*R}.  (0, 1*R) is the
// reference's encoding (constants.cpp:264-292); (0, 0) is its zero padding
// (ecbn128_device.cu:1836-1841).  Neither is a curve point since b != 0, 1.
template <class F> __device__ __forceinline__ bool affine_is_inf(const Affine<F> &p) {
  return p.x.is_zero() && (p.y.is_zero() || p.y.is_one());
}

}  // namespace csk
