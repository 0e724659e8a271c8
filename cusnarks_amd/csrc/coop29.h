// Wave-cooperative point formulas on 29-bit limbs for the window combination
// (the serial Horner chain that closes a multiexp).  G2 values live in LDS
// slots run by a small interpreter (below); G1 points are held in every
// lane's registers (end of file), which is faster for G1: its sums and
// differences need no round of their own.
//
// Every value of a formula lives in an LDS slot (one 9-limb Fq29 component).
// A formula is a program of rounds; in a round, lane i executes operation i
// (a product, a one-reduction sum of two or four products, or a biased
// add/subtract/conditional subtract), reading its operands from slots and
// writing its result to a slot, and the round ends with a barrier.  The
// independent products of a step thus run on different lanes at once, and
// no lane holds more than one operation's operands in registers.  The ranges
// are those of the G2 bucket loop (field29.h): X < 4q per component, Y, ZZ,
// ZZZ < 2q; each operation's bound is noted in its program.  Lanes of one
// round that run different operation kinds take turns (the wave walks each
// kind's path), so a round costs the sum over its kinds: a round of
// products keeps to one kind.
#pragma once
#include "field29.h"

namespace csk {

enum : uint8_t {
  C29_MUL,    // d = a0 a1
  C29_DOT,    // d = a0 a1 + a2 a3
  C29_DOT4,   // d = a0 a1 + a2 a3 + a4 a5 + a6 a7
  C29_ADD1,   // d = a0 + a1
  C29_ADD2,   // d = a0 + 2 a1
  C29_SUB2,   // d = a0 + 2q - a1
  C29_SUB4,   // d = a0 + 4q - a1
  C29_SUB6,   // d = a0 + 6q - a1
  C29_SUB8,   // d = a0 + 8q - a1
  C29_CSUB4,  // d = a0 - 4q if a0 >= 4q
  C29_COPY,   // d = a0
  // chains of the above in one lane, each exactly the composition named, so
  // a step needs no round per link
  C29_SUB4D,  // d = SUB4(a0, ADD1(a1, a1))
  C29_SUB6T,  // d = SUB6(a0, ADD2(a1, a1))
  C29_NSUB,   // d = SUB4(0, SUB2(a0, a1))
  C29_X3D,    // x = CSUB4(SUB4(a0, ADD1(a1, a1))); d = x        doubling: X3 from MM, S
  C29_TD,     // d = SUB4(a1, x)                                        T = S - X3
  C29_NTD,    // d = SUB6(0, SUB4(a1, x))                               -T
  C29_X3A,    // x = CSUB4(SUB6(a0, ADD2(a1, a2))); d = x        addition: X3 from RR, PPP, Qv
  C29_TA,     // d = SUB4(a2, x)                                        T = Qv - X3
  C29_NTA,    // d = SUB6(0, SUB4(a2, x))                               -T
  // a two-term sum and what its lane derives from it (a second slot in a4)
  C29_DOTN,   // d = DOT(a0..a3); a4 = SUB2(0, d)
  C29_DOTM,   // d = ADD2(x, x), x = DOT(a0..a3)
  C29_DOTMN,  // d = ADD2(x, x), x = DOT(a0..a3); a4 = SUB6(0, d)
};
struct C29Op {
  uint8_t k, d, a[8];
};

// redc(a0 b0 + a1 b1 + a2 b2 + a3 b3), normalised operands, latency-ordered
// as mul29_lat
__device__ __forceinline__ Fq29 dot4_29_lat(const Fq29 &a0, const Fq29 &b0, const Fq29 &a1, const Fq29 &b1,
                                            const Fq29 &a2, const Fq29 &b2, const Fq29 &a3, const Fq29 &b3) {
  uint64_t t[17] = {};
  cols29_mad(t, a0, b0);
  cols29_mad(t, a1, b1);
  cols29_mad(t, a2, b2);
  cols29_mad(t, a3, b3);
  return cols29_redc(t);
}

__device__ __forceinline__ Fq29 c29_ld(const uint32_t *sh, uint32_t s) {
  Fq29 r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.l[i] = sh[9 * s + i];
  return r;
}
__device__ __forceinline__ void c29_st(uint32_t *sh, uint32_t s, const Fq29 &v) {
#pragma unroll
  for (int i = 0; i < 9; i++) sh[9 * s + i] = v.l[i];
}

__device__ __forceinline__ void c29_exec(uint32_t *sh, const C29Op op) {
  const Fq29 a = c29_ld(sh, op.a[0]);
  Fq29 r;
  switch (op.k) {
    case C29_MUL: r = mul29_lat(a, c29_ld(sh, op.a[1])); break;
    case C29_DOT:
    case C29_DOTN:
    case C29_DOTM:
    case C29_DOTMN:  // one path for the four
      r = dot29_lat(a, c29_ld(sh, op.a[1]), c29_ld(sh, op.a[2]), c29_ld(sh, op.a[3]));
      if (op.k == C29_DOTM || op.k == C29_DOTMN) r = add29<2>(r, r);
      if (op.k == C29_DOTN) c29_st(sh, op.a[4], sub29<2>(Fq29{}, r));
      if (op.k == C29_DOTMN) c29_st(sh, op.a[4], sub29<6>(Fq29{}, r));
      break;
    case C29_DOT4:
      r = dot4_29_lat(a, c29_ld(sh, op.a[1]), c29_ld(sh, op.a[2]), c29_ld(sh, op.a[3]), c29_ld(sh, op.a[4]),
                      c29_ld(sh, op.a[5]), c29_ld(sh, op.a[6]), c29_ld(sh, op.a[7]));
      break;
    case C29_ADD1: r = add29<1>(a, c29_ld(sh, op.a[1])); break;
    case C29_ADD2: r = add29<2>(a, c29_ld(sh, op.a[1])); break;
    case C29_SUB2: r = sub29<2>(a, c29_ld(sh, op.a[1])); break;
    case C29_SUB4: r = sub29<4>(a, c29_ld(sh, op.a[1])); break;
    case C29_SUB6: r = sub29<6>(a, c29_ld(sh, op.a[1])); break;
    case C29_SUB8: r = sub29<8>(a, c29_ld(sh, op.a[1])); break;
    case C29_CSUB4: r = csub29(a, Q29::p4); break;
    case C29_SUB4D: {
      const Fq29 b = c29_ld(sh, op.a[1]);
      r = sub29<4>(a, add29<1>(b, b));
      break;
    }
    case C29_SUB6T: {
      const Fq29 b = c29_ld(sh, op.a[1]);
      r = sub29<6>(a, add29<2>(b, b));
      break;
    }
    case C29_NSUB: r = sub29<4>(Fq29{}, sub29<2>(a, c29_ld(sh, op.a[1]))); break;
    case C29_X3D:
    case C29_TD:
    case C29_NTD: {  // one path for the three: the lanes of a round share it
      const Fq29 sv = c29_ld(sh, op.a[1]);
      const Fq29 x = csub29(sub29<4>(a, add29<1>(sv, sv)), Q29::p4);
      r = op.k == C29_X3D ? x : sub29<4>(sv, x);
      if (op.k == C29_NTD) r = sub29<6>(Fq29{}, r);
      break;
    }
    case C29_X3A:
    case C29_TA:
    case C29_NTA: {
      const Fq29 qv = c29_ld(sh, op.a[2]);
      const Fq29 x = csub29(sub29<6>(a, add29<2>(c29_ld(sh, op.a[1]), qv)), Q29::p4);
      r = op.k == C29_X3A ? x : sub29<4>(qv, x);
      if (op.k == C29_NTA) r = sub29<6>(Fq29{}, r);
      break;
    }
    default: r = a; break;
  }
  c29_st(sh, op.d, r);
}

// rounds[r] operations per round, taken in order from ops
__device__ __forceinline__ void c29_run(uint32_t *sh, const C29Op *ops, const uint8_t *rounds, int nrounds) {
  uint32_t off = 0;
  for (int r = 0; r < nrounds; r++) {
    const uint32_t n = rounds[r];
    if (threadIdx.x < n) c29_exec(sh, ops[off + threadIdx.x]);
    __syncthreads();
    off += n;
  }
}

namespace {  // programs (one copy per translation unit)

// ---- G2: acc components X0 X1 Y0 Y1 ZZ0 ZZ1 ZZZ0 ZZZ1 in 0-7, addend in
// 8-15, zero in 16.  An Fq2 product is two reductions: c0 = a0 b0 + a1 (Kq -
// b1), c1 = a0 b1 + a1 b0 (squares too, with the negated a1).
enum : uint8_t { H_X0, H_X1, H_Y0, H_Y1, H_ZZ0, H_ZZ1, H_ZZZ0, H_ZZZ1,
                 H_QX0, H_QX1, H_QY0, H_QY1, H_QZZ0, H_QZZ1, H_QZZZ0, H_QZZZ1, H_ZERO };

// 2 acc: U = 2Y < 4q; V = U^2, xx = X^2 (32 q^2); M = 3 xx < 6q; M^2 (72 q^2);
// X3 = M^2 - 2S < 6q, then < 4q; Y3 = M (S - X3) - W Y (80 q^2)
enum : uint8_t { D2_U0 = 17, D2_U1, D2_NY0, D2_NY1, D2_NZZ1, D2_NZZZ1, D2_NX1, D2_NU1, D2_V0, D2_V1, D2_XX0, D2_XX1,
                 D2_M0, D2_M1, D2_NV1, D2_NM1, D2_W0, D2_W1, D2_S0, D2_S1, D2_MM0, D2_MM1, D2_ZZ30, D2_ZZ31,
                 D2_S20, D2_S21, D2_X30, D2_X31, D2_T0, D2_T1, D2_NT1, D2_Y30, D2_Y31, D2_ZZZ30, D2_ZZZ31,
                 D2_OY1, D2_OZZZ0, D2_OZZZ1 };
// The last rounds write the accumulator slots in place (X3 and ZZ3 once no
// later round reads the old X and ZZ, Y3 and ZZZ3 from copies of the old
// Y1, ZZZ0, ZZZ1 taken in the first round), so no round of copies follows.
__constant__ C29Op kG2Dbl[] = {
    {C29_ADD1, D2_U0, {H_Y0, H_Y0}}, {C29_ADD1, D2_U1, {H_Y1, H_Y1}}, {C29_SUB2, D2_NY0, {H_ZERO, H_Y0}},
    {C29_SUB2, D2_NY1, {H_ZERO, H_Y1}}, {C29_SUB2, D2_NZZ1, {H_ZERO, H_ZZ1}}, {C29_SUB2, D2_NZZZ1, {H_ZERO, H_ZZZ1}},
    {C29_SUB4, D2_NX1, {H_ZERO, H_X1}}, {C29_SUB4D, D2_NU1, {H_ZERO, H_Y1}},  // 4q - U1
    {C29_COPY, D2_OY1, {H_Y1}}, {C29_COPY, D2_OZZZ0, {H_ZZZ0}}, {C29_COPY, D2_OZZZ1, {H_ZZZ1}},
    // V = U^2 (and 2q - V1); M = 3 X^2 (and 6q - M1)
    {C29_DOT, D2_V0, {D2_U0, D2_U0, D2_U1, D2_NU1}}, {C29_DOTN, D2_V1, {D2_U0, D2_U1, D2_U1, D2_U0, D2_NV1}},
    {C29_DOTM, D2_M0, {H_X0, H_X0, H_X1, D2_NX1}}, {C29_DOTMN, D2_M1, {H_X0, H_X1, H_X1, H_X0, D2_NM1}},
    {C29_DOT, D2_W0, {D2_U0, D2_V0, D2_U1, D2_NV1}}, {C29_DOT, D2_W1, {D2_U0, D2_V1, D2_U1, D2_V0}},
    {C29_DOT, D2_S0, {H_X0, D2_V0, H_X1, D2_NV1}},   {C29_DOT, D2_S1, {H_X0, D2_V1, H_X1, D2_V0}},
    {C29_DOT, D2_MM0, {D2_M0, D2_M0, D2_M1, D2_NM1}}, {C29_DOT, D2_MM1, {D2_M0, D2_M1, D2_M1, D2_M0}},
    {C29_DOT, D2_ZZ30, {D2_V0, H_ZZ0, D2_V1, D2_NZZ1}}, {C29_DOT, D2_ZZ31, {D2_V0, H_ZZ1, D2_V1, H_ZZ0}},
    // X3 = MM - 2S < 6q, then < 4q; T = S - X3 < 6q; 6q - T1
    {C29_X3D, H_X0, {D2_MM0, D2_S0}}, {C29_X3D, H_X1, {D2_MM1, D2_S1}},
    {C29_TD, D2_T0, {D2_MM0, D2_S0}},   {C29_TD, D2_T1, {D2_MM1, D2_S1}},
    {C29_NTD, D2_NT1, {D2_MM1, D2_S1}}, {C29_COPY, H_ZZ0, {D2_ZZ30}}, {C29_COPY, H_ZZ1, {D2_ZZ31}},
    // ZZZ3 as four-term sums with two zero terms: the same column sums as
    // C29_DOT, so the same result, and the round runs one operation kind
    {C29_DOT4, H_Y0, {D2_M0, D2_T0, D2_M1, D2_NT1, D2_W0, D2_NY0, D2_W1, D2_OY1}},
    {C29_DOT4, H_Y1, {D2_M0, D2_T1, D2_M1, D2_T0, D2_W0, D2_NY1, D2_W1, D2_NY0}},
    {C29_DOT4, H_ZZZ0, {D2_W0, D2_OZZZ0, D2_W1, D2_NZZZ1, H_ZERO, H_ZERO, H_ZERO, H_ZERO}},
    {C29_DOT4, H_ZZZ1, {D2_W0, D2_OZZZ1, D2_W1, D2_OZZZ0, H_ZERO, H_ZERO, H_ZERO, H_ZERO}},
};
__constant__ uint8_t kG2DblR[] = {11, 4, 8, 7, 4};

// acc + q, part 1: U1 = X qZZ, U2 = qX ZZ, S1 = Y qZZZ, S2 = qY ZZZ, ZZ qZZ,
// ZZZ qZZZ (16 q^2 each); P, R < 4q; PP, RR (32 q^2)
enum : uint8_t { E_NQZZ1 = 17, E_NQZZZ1, E_NZZ1, E_NZZZ1, E_U10, E_U11, E_U20, E_U21, E_S10, E_S11, E_S20, E_S21,
                 E_Z20, E_Z21, E_Z30, E_Z31, E_P0, E_P1, E_R0, E_R1, E_NP1, E_NR1, E_PP0, E_PP1, E_RR0, E_RR1,
                 E_NPP1, E_PPP0, E_PPP1, E_QV0, E_QV1, E_ZZ30, E_ZZ31, E_NPPP1, E_NP0, E_T10, E_T11, E_X30, E_X31,
                 E_T0, E_T1, E_NT1, E_Y30, E_Y31, E_ZZZ30, E_ZZZ31, E_NSLOT };
__constant__ C29Op kG2Add1[] = {
    {C29_SUB2, E_NQZZ1, {H_ZERO, H_QZZ1}}, {C29_SUB2, E_NQZZZ1, {H_ZERO, H_QZZZ1}},
    {C29_SUB2, E_NZZ1, {H_ZERO, H_ZZ1}},   {C29_SUB2, E_NZZZ1, {H_ZERO, H_ZZZ1}},
    {C29_DOT, E_U10, {H_X0, H_QZZ0, H_X1, E_NQZZ1}},     {C29_DOT, E_U11, {H_X0, H_QZZ1, H_X1, H_QZZ0}},
    {C29_DOT, E_U20, {H_QX0, H_ZZ0, H_QX1, E_NZZ1}},     {C29_DOT, E_U21, {H_QX0, H_ZZ1, H_QX1, H_ZZ0}},
    {C29_DOT, E_S10, {H_Y0, H_QZZZ0, H_Y1, E_NQZZZ1}},   {C29_DOT, E_S11, {H_Y0, H_QZZZ1, H_Y1, H_QZZZ0}},
    {C29_DOT, E_S20, {H_QY0, H_ZZZ0, H_QY1, E_NZZZ1}},   {C29_DOT, E_S21, {H_QY0, H_ZZZ1, H_QY1, H_ZZZ0}},
    {C29_DOT, E_Z20, {H_ZZ0, H_QZZ0, H_ZZ1, E_NQZZ1}},   {C29_DOT, E_Z21, {H_ZZ0, H_QZZ1, H_ZZ1, H_QZZ0}},
    {C29_DOT, E_Z30, {H_ZZZ0, H_QZZZ0, H_ZZZ1, E_NQZZZ1}}, {C29_DOT, E_Z31, {H_ZZZ0, H_QZZZ1, H_ZZZ1, H_QZZZ0}},
    {C29_SUB2, E_P0, {E_U20, E_U10}}, {C29_SUB2, E_P1, {E_U21, E_U11}},
    {C29_SUB2, E_R0, {E_S20, E_S10}}, {C29_SUB2, E_R1, {E_S21, E_S11}},
    {C29_NSUB, E_NP1, {E_U21, E_U11}}, {C29_NSUB, E_NR1, {E_S21, E_S11}},  // 4q - P1, 4q - R1
    {C29_DOT, E_PP0, {E_P0, E_P0, E_P1, E_NP1}}, {C29_DOT, E_PP1, {E_P0, E_P1, E_P1, E_P0}},
    {C29_DOT, E_RR0, {E_R0, E_R0, E_R1, E_NR1}}, {C29_DOT, E_RR1, {E_R0, E_R1, E_R1, E_R0}},
};
__constant__ uint8_t kG2Add1R[] = {4, 12, 6, 4};
// part 2: PPP, Qv, ZZ3 (16 q^2); X3 = RR - PPP - 2Qv < 8q, then < 4q;
// T = Qv - X3 < 6q; Y3 = R T - S1 PPP (56 q^2); ZZZ3
__constant__ C29Op kG2Add2[] = {
    {C29_SUB2, E_NPP1, {H_ZERO, E_PP1}},
    {C29_DOT, E_PPP0, {E_P0, E_PP0, E_P1, E_NPP1}},  {C29_DOT, E_PPP1, {E_P0, E_PP1, E_P1, E_PP0}},
    {C29_DOT, E_QV0, {E_U10, E_PP0, E_U11, E_NPP1}}, {C29_DOT, E_QV1, {E_U10, E_PP1, E_U11, E_PP0}},
    {C29_DOT, E_ZZ30, {E_Z20, E_PP0, E_Z21, E_NPP1}}, {C29_DOT, E_ZZ31, {E_Z20, E_PP1, E_Z21, E_PP0}},
    {C29_SUB2, E_NPPP1, {H_ZERO, E_PPP1}}, {C29_SUB2, E_NP0, {H_ZERO, E_PPP0}},
    {C29_X3A, H_X0, {E_RR0, E_PPP0, E_QV0}}, {C29_X3A, H_X1, {E_RR1, E_PPP1, E_QV1}},  // in place, as in kG2Dbl
    {C29_TA, E_T0, {E_RR0, E_PPP0, E_QV0}},   {C29_TA, E_T1, {E_RR1, E_PPP1, E_QV1}},
    {C29_NTA, E_NT1, {E_RR1, E_PPP1, E_QV1}}, {C29_COPY, H_ZZ0, {E_ZZ30}}, {C29_COPY, H_ZZ1, {E_ZZ31}},
    {C29_DOT4, H_Y0, {E_R0, E_T0, E_R1, E_NT1, E_S10, E_NP0, E_S11, E_PPP1}},
    {C29_DOT4, H_Y1, {E_R0, E_T1, E_R1, E_T0, E_S10, E_NPPP1, E_S11, E_NP0}},
    {C29_DOT4, H_ZZZ0, {E_Z30, E_PPP0, E_Z31, E_NPPP1, H_ZERO, H_ZERO, H_ZERO, H_ZERO}},
    {C29_DOT4, H_ZZZ1, {E_Z30, E_PPP1, E_Z31, E_PPP0, H_ZERO, H_ZERO, H_ZERO, H_ZERO}},
};
__constant__ uint8_t kG2Add2R[] = {1, 6, 9, 4};

}  // namespace

// Programs and slot counts for k_final.  The programs are copied into LDS
// once per workgroup (load): each lane fetches a different operation every
// round, which from constant memory is a vector load and its full latency.
template <class F> struct Coop29;
template <> struct Coop29<Fq2> {
  static constexpr int NC = 8;
  static constexpr int kSlots = E_NSLOT > D2_OZZZ1 + 1 ? E_NSLOT : D2_OZZZ1 + 1;
  static constexpr int kDbl = sizeof(kG2Dbl) / sizeof(C29Op), kAdd1 = sizeof(kG2Add1) / sizeof(C29Op),
                       kAdd2 = sizeof(kG2Add2) / sizeof(C29Op), kOps = kDbl + kAdd1 + kAdd2;
  __device__ static void load(C29Op *prog) {
    for (int i = threadIdx.x; i < kOps; i += blockDim.x)
      prog[i] = i < kDbl ? kG2Dbl[i] : i < kDbl + kAdd1 ? kG2Add1[i - kDbl] : kG2Add2[i - kDbl - kAdd1];
    __syncthreads();
  }
  __device__ static void dbl(uint32_t *sh, const C29Op *prog) { c29_run(sh, prog, kG2DblR, sizeof(kG2DblR)); }
  __device__ static void add1(uint32_t *sh, const C29Op *prog) {
    c29_run(sh, prog + kDbl, kG2Add1R, sizeof(kG2Add1R));
  }
  __device__ static void add2(uint32_t *sh, const C29Op *prog) {
    c29_run(sh, prog + kDbl + kAdd1, kG2Add2R, sizeof(kG2Add2R));
  }
  static constexpr uint8_t ZZ[2] = {H_ZZ0, H_ZZ1}, QZZ[2] = {H_QZZ0, H_QZZ1}, PP[2] = {E_PP0, E_PP1},
                           RR[2] = {E_RR0, E_RR1};
};

// zero mod q for both (or the one) component slots
__device__ __forceinline__ bool c29_is_zero(const uint32_t *sh, const uint8_t (&s)[2]) {
  return is_zero29(c29_ld(sh, s[0])) && is_zero29(c29_ld(sh, s[1]));
}

// ---- G1: register-resident cooperative formulas.  Every lane holds the
// point as 29-bit limbs; the products of a round (products, or one-reduction
// sums of two products) run on lanes 0..K-1 and are exchanged through LDS;
// sums and differences are computed by every lane in registers, without a
// round of their own.  Operations and ranges as xdbl29 / xadd29.
template <int K>
__device__ __forceinline__ void coop29_mul(Fq29 *sh, const Fq29 &a, const Fq29 &b, Fq29 (&out)[K]) {
  const Fq29 r = mul29_lat(a, b);
  if (threadIdx.x < (uint32_t)K) sh[threadIdx.x] = r;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) out[k] = sh[k];
  __syncthreads();
}
template <int K>
__device__ __forceinline__ void coop29_dot(Fq29 *sh, const Fq29 &a, const Fq29 &b, const Fq29 &c, const Fq29 &d,
                                           Fq29 (&out)[K]) {
  const Fq29 r = dot29_lat(a, b, c, d);
  if (threadIdx.x < (uint32_t)K) sh[threadIdx.x] = r;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) out[k] = sh[k];
  __syncthreads();
}

// lane sel takes operand pair k (selects on named values: an indexed operand
// array would be placed in scratch memory)
#define C29_PICK(k, A, B) \
  if (sel == (k)) {       \
    a = (A);              \
    b = (B);              \
  }

// p = 2p (p not at infinity)
__device__ __forceinline__ void xdbl29_coop(Fq29 *sh, XYZZ29 &p) {
  const uint32_t sel = threadIdx.x;
  const Fq29 U = add29<1>(p.Y, p.Y);
  Fq29 a = U, b = U, r1[2], r2[4], r3[2];
  C29_PICK(1, p.X, p.X);
  coop29_mul<2>(sh, a, b, r1);  // V = U^2, xx = X^2
  const Fq29 V = r1[0], M = add29<2>(r1[1], r1[1]);
  a = U;
  b = V;
  C29_PICK(1, p.X, V);
  C29_PICK(2, M, M);
  C29_PICK(3, V, p.ZZ);
  coop29_mul<4>(sh, a, b, r2);  // W, S, M^2, ZZ3
  const Fq29 W = r2[0], S = r2[1];
  const Fq29 X3 = sub29<4>(r2[2], add29<1>(S, S));
  Fq29 c = W, d = sub29<2>(Fq29{}, p.Y);
  a = M;
  b = sub29<6>(S, X3);
  if (sel == 1) {
    a = W;
    b = p.ZZZ;
    c = Fq29{};
    d = Fq29{};
  }
  coop29_dot<2>(sh, a, b, c, d, r3);  // Y3 = M (S - X3) - W Y, ZZZ3 = W ZZZ
  p = {X3, r3[0], r2[3], r3[1]};
}

// p += q (neither at infinity); sets inf when q == -p
__device__ __forceinline__ void xadd29_coop(Fq29 *sh, XYZZ29 &p, const XYZZ29 &q, bool &inf) {
  const uint32_t sel = threadIdx.x;
  Fq29 a = p.X, b = q.ZZ, r1[6], r2[2], r3[3], r4[2];
  C29_PICK(1, q.X, p.ZZ);
  C29_PICK(2, p.Y, q.ZZZ);
  C29_PICK(3, q.Y, p.ZZZ);
  C29_PICK(4, p.ZZ, q.ZZ);
  C29_PICK(5, p.ZZZ, q.ZZZ);
  coop29_mul<6>(sh, a, b, r1);  // U1, U2, S1, S2, ZZ1 ZZ2, ZZZ1 ZZZ2
  const Fq29 U1 = r1[0], S1 = r1[2];
  const Fq29 P = sub29<2>(r1[1], U1), R = sub29<2>(r1[3], S1);  // < 4q
  a = P;
  b = P;
  C29_PICK(1, R, R);
  coop29_mul<2>(sh, a, b, r2);  // PP, RR
  if (is_zero29(r2[0])) {
    if (is_zero29(r2[1])) xdbl29_coop(sh, p);  // q == p
    else inf = true;                            // q == -p
    return;
  }
  const Fq29 PP = r2[0];
  a = P;
  b = PP;
  C29_PICK(1, U1, PP);
  C29_PICK(2, r1[4], PP);
  coop29_mul<3>(sh, a, b, r3);  // PPP, Qv, ZZ3
  const Fq29 PPP = r3[0], Qv = r3[1];
  const Fq29 X3 = sub29<6>(r2[1], add29<2>(PPP, Qv));  // < 8q
  Fq29 c = S1, d = sub29<2>(Fq29{}, PPP);
  a = R;
  b = sub29<8>(Qv, X3);
  if (sel == 1) {
    a = r1[5];
    b = PPP;
    c = Fq29{};
    d = Fq29{};
  }
  coop29_dot<2>(sh, a, b, c, d, r4);  // Y3 = R (Qv - X3) - S1 PPP, ZZZ3
  p = {X3, r4[0], r3[2], r4[1]};
}
#undef C29_PICK

}  // namespace csk
