/*
 * CPU oracle for the cusnarks_amd hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU algorithms (cusnarks
 * src/cuda/ff.cpp, ec.cpp, ntt.cpp, mpoly.cpp).  It is compiled into
 * oracle/liboracle.so and is used solely by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg, as the checker and the timed CPU
 * baseline.  The product library (cusnarks_amd/libcsk.so) never links or
 * calls it.
 *
 * Data layout follows the reference: every field element is 8 x u32
 * little-endian limbs (Montgomery, R = 2^256); pidx 0 = base field q
 * ("MOD_GROUP"), pidx 1 = scalar field r ("MOD_FIELD").
 */
#ifndef CSK_ORACLE_H
#define CSK_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* field (ff.cpp) */
void or_montmult(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t pidx);
void or_montsquare(uint32_t *z, const uint32_t *x, uint32_t pidx);
void or_addm(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t pidx);
void or_subm(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t pidx);
void or_montinv(uint32_t *z, const uint32_t *x, uint32_t pidx);
void or_to_montgomeryN(uint32_t *z, const uint32_t *x, uint32_t n, uint32_t pidx);
void or_from_montgomeryN(uint32_t *z, const uint32_t *x, uint32_t n, uint32_t pidx);
void or_montmultN(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t n, uint32_t pidx);
void or_montmult_ext(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t pidx);

/* elliptic curve (ec.cpp) -- Jacobian (X,Y,Z), G2 coordinates are Fq2 */
void or_ec_jacadd(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t ec2);
void or_ec_jacaddmixed(uint32_t *z, const uint32_t *x_aff, const uint32_t *y, uint32_t ec2);
void or_ec_jacdouble(uint32_t *z, const uint32_t *x, uint32_t ec2);
void or_ec_jac2aff(uint32_t *y, const uint32_t *x, uint32_t n, uint32_t ec2, uint32_t strip_last);
/* Straus multi-exponentiation server (ec_jacreduce_server_h): out = affine
 * Montgomery (X, Y) of sum k_i P_i.  max_threads = 0 -> min(nprocs, 32). */
int or_ec_jacreduce(uint32_t *out_aff, const uint32_t *scl, const uint32_t *pts,
                    uint32_t n, uint32_t order, uint32_t ec2, uint32_t max_threads,
                    uint32_t *threads_used);
/* fixed-base k*G for generating test inputs (normal-form scalars) */
void or_ec_scmul_aff(uint32_t *out_aff, const uint32_t *scl, const uint32_t *pt_aff,
                     uint32_t n, uint32_t ec2);

/* NTT (ntt.cpp) over Fr, natural order in and out, Montgomery data */
void or_roots(uint32_t *roots, uint32_t log2n);
void or_ntt(uint32_t *a, const uint32_t *roots, uint32_t log2n, uint32_t rstride);
void or_intt(uint32_t *a, const uint32_t *roots, uint32_t log2n, uint32_t rstride, uint32_t fmt);
/* H step: out (2m words x 8) = INTT_2m of [A.B on <w_m> | A'.B' on the odd coset],
 * normal form; roots = 2m-th roots (Montgomery).  */
void or_interpolandmul(uint32_t *out, const uint32_t *pA, const uint32_t *pB,
                       const uint32_t *roots, uint32_t log2m, uint32_t nthreads);
/* the same, computed as the reference's threaded server does (4-step
 * transforms split over nthreads threads); results identical */
void or_interpolandmul_mt(uint32_t *out, const uint32_t *pA, const uint32_t *pB,
                          const uint32_t *roots, uint32_t log2m, uint32_t nthreads);
/* closed-form checks: Horner value (normal-form coefficients, x Montgomery,
 * normal result) and the barycentric value at z of the interpolant of m
 * Montgomery evaluations on <w_m> (Montgomery result; -1 if z is in the domain) */
void or_poly_eval(uint32_t *out, const uint32_t *coeffs, uint64_t n, const uint32_t *x_mont);
int or_lagrange_eval(uint32_t *out, const uint32_t *evals, uint32_t log2m, const uint32_t *z);
/* sparse QAP evaluation: CSR rows -> out[row] = sum toMont(w[col]) * val */
void or_mpoly_eval_csr(uint32_t *out, const uint64_t *rowptr, const uint32_t *col,
                       const uint32_t *val, const uint32_t *wit, uint32_t nrows);

#ifdef __cplusplus
}
#endif
#endif
