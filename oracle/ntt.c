/*
 * Oracle NTT / H-polynomial / QAP evaluation -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates, over Fr (pidx 1):
 *   computeRoots_h   cusnarks src/cuda/ntt.cpp:1181-1195 (roots[k] = w^k, Montgomery,
 *                    w = proots_field[log2n] = 5^((r-1)/2^log2n), constants.cpp:323-352)
 *   ntt_h / _ntt_h   ntt.cpp:392-463 (bit-reverse reorder, radix-2 DIT, table index
 *                    (rstride*k) & (rstride*n - 1))
 *   intt_h           ntt.cpp:583-603 (same with direction -1, then x IScaler[log2n]:
 *                    fmt 1 -> N^-1 R (Montgomery out), fmt 0 -> N^-1 (normal out))
 *   interpol_odd / ntt_interpolandmul_parallel_h ntt.cpp:636-670, 1072-1105:
 *     M[2i] = A[i]B[i]; A,B <- NTT(INTT(.) * w_2m^i); M[2i+1] = A[i]B[i];
 *     M <- INTT_2m(M) in normal form.
 *   ntt_interpolandmul_server_h ntt.cpp:672-765 (threaded, 4-step transforms of
 *                    ntt_parallel_h 274-316): or_interpolandmul_mt, the timed CPU
 *                    baseline of the H step; pinned bit for bit against
 *                    or_interpolandmul by tests/test_oracle.py
 *   mpoly_eval_h     mpoly.cpp:125-168 (pX[row] += coeff * toMont(w[wire]))
 */
#include <stdlib.h>
#include <string.h>
#include "csk_oracle.h"
#include "oracle_internal.h"

static const uint32_t ROOT28_MONT[8] = {2161196444, 1668182869, 608567254, 2720789364,
                                        514982872,  1447373504, 694417895, 409005972};

void or_roots(uint32_t *roots, uint32_t log2n)
{
    /* w_n = w_{2^28}^(2^(28 - log2n)) */
    uint32_t w[8];
    memcpy(w, ROOT28_MONT, 32);
    for (uint32_t i = log2n; i < 28; i++) or_montmult(w, w, w, 1);
    size_t n = (size_t)1 << log2n;
    memcpy(roots, OR_ONE[1], 32);
    for (size_t i = 1; i < n; i++) or_montmult(&roots[8 * i], &roots[8 * (i - 1)], w, 1);
}

static uint32_t bitrev(uint32_t x, uint32_t bits)
{
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

static void reorder(uint32_t *a, uint32_t log2n)
{
    uint32_t n = 1u << log2n;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t j = bitrev(i, log2n);
        if (j > i) {
            uint32_t t[8];
            memcpy(t, &a[8 * (size_t)i], 32);
            memcpy(&a[8 * (size_t)i], &a[8 * (size_t)j], 32);
            memcpy(&a[8 * (size_t)j], t, 32);
        }
    }
}

static void dit(uint32_t *a, const uint32_t *roots, uint32_t log2n, uint64_t rstride, int dir)
{
    int64_t n = (int64_t)1 << log2n;
    uint64_t mask = rstride * (uint64_t)n - 1;
    uint32_t right[8];
    for (int64_t size = 2; size <= n; size *= 2) {
        int64_t half = size >> 1, step = dir * (n / size);
        for (int64_t i = 0; i < n; i += size) {
            int64_t k = 0;
            for (int64_t j = i; j < i + half; j++) {
                uint64_t ridx = ((uint64_t)((int64_t)rstride * k)) & mask;
                uint32_t *L = &a[8 * (j + half)], *J = &a[8 * j];
                or_montmult(right, L, &roots[8 * ridx], 1);
                or_subm(L, J, right, 1);
                or_addm(J, J, right, 1);
                k += step;
            }
        }
    }
}

void or_ntt(uint32_t *a, const uint32_t *roots, uint32_t log2n, uint32_t rstride)
{
    reorder(a, log2n);
    dit(a, roots, log2n, rstride, 1);
}

static void scaler(uint32_t *s, uint32_t log2n, uint32_t fmt)
{
    /* N^-1 in Montgomery form, then optionally leave the Montgomery domain */
    uint32_t nm[8] = {0}, one[8] = {1, 0, 0, 0, 0, 0, 0, 0};
    nm[log2n >> 5] = 1u << (log2n & 31);
    or_to_montgomeryN(nm, nm, 1, 1);
    or_montinv(s, nm, 1);
    if (fmt == 0) or_montmult(s, s, one, 1);
}

void or_intt(uint32_t *a, const uint32_t *roots, uint32_t log2n, uint32_t rstride, uint32_t fmt)
{
    uint32_t s[8];
    reorder(a, log2n);
    dit(a, roots, log2n, rstride, -1);
    scaler(s, log2n, fmt);
    size_t n = (size_t)1 << log2n;
    for (size_t i = 0; i < n; i++) or_montmult(&a[8 * i], &a[8 * i], s, 1);
}

static void interpol_odd(uint32_t *a, const uint32_t *roots, uint32_t log2m)
{
    size_t m = (size_t)1 << log2m;
    or_intt(a, roots, log2m, 2, 1);
    for (size_t i = 0; i < m; i++) or_montmult(&a[8 * i], &a[8 * i], &roots[8 * i], 1);
    or_ntt(a, roots, log2m, 2);
}

/* 4-step transform of N = 2^L elements, the threaded form the reference's
 * ntt_interpolandmul_server_h uses (ntt_parallel_h / ntt_parallel_T_h,
 * ntt.cpp:274-316; scaled inverse intt_parallel_h 361-374): the array read as
 * a 2^nr x 2^nc matrix (row j, column i = element j 2^nc + i), transforms of
 * every column (roots stride rstride << nc), the twiddles w^(dir k1 i), row
 * transforms (stride rstride << nr), then the transpose, so the output is in
 * natural order.  Columns, twiddles and rows are split over the threads. */
static void ntt4(uint32_t *a, const uint32_t *roots, uint32_t L, uint64_t rstride, int dir,
                 uint32_t *tmp, int nthr)
{
    const uint32_t nr = L / 2, nc = L - nr;
    const size_t R = (size_t)1 << nr, C = (size_t)1 << nc, N = R * C;
    const uint64_t mask = rstride * (uint64_t)N - 1;
#pragma omp parallel num_threads(nthr)
    {
        uint32_t *col = (uint32_t *)malloc(R * 32);
#pragma omp for schedule(static)
        for (size_t i = 0; i < C; i++) {
            for (size_t j = 0; j < R; j++) memcpy(&col[8 * j], &a[8 * (j * C + i)], 32);
            reorder(col, nr);
            dit(col, roots, nr, rstride << nc, dir);
            for (size_t j = 0; j < R; j++) memcpy(&a[8 * (j * C + i)], &col[8 * j], 32);
        }
        free(col);
#pragma omp for schedule(static)
        for (size_t e = 0; e < N; e++) {
            uint64_t ridx = ((uint64_t)((int64_t)rstride * (int64_t)(e >> nc) * (int64_t)(e & (C - 1)) * dir)) & mask;
            or_montmult(&a[8 * e], &a[8 * e], &roots[8 * ridx], 1);
        }
#pragma omp for schedule(static)
        for (size_t j = 0; j < R; j++) {
            reorder(&a[8 * j * C], nc);
            dit(&a[8 * j * C], roots, nc, rstride << nr, dir);
        }
#pragma omp for schedule(static)
        for (size_t j = 0; j < R; j++)
            for (size_t i = 0; i < C; i++) memcpy(&tmp[8 * (i * R + j)], &a[8 * (j * C + i)], 32);
#pragma omp for schedule(static)
        for (size_t e = 0; e < N; e++) memcpy(&a[8 * e], &tmp[8 * e], 32);
    }
}

static void intt4(uint32_t *a, const uint32_t *roots, uint32_t L, uint64_t rstride, uint32_t fmt,
                  uint32_t *tmp, int nthr)
{
    uint32_t s[8];
    ntt4(a, roots, L, rstride, -1, tmp, nthr);
    scaler(s, L, fmt);
    size_t n = (size_t)1 << L;
#pragma omp parallel for num_threads(nthr) schedule(static)
    for (size_t i = 0; i < n; i++) or_montmult(&a[8 * i], &a[8 * i], s, 1);
}

/* The H step as ntt_interpolandmul_server_h computes it with nthreads
 * threads (ntt.cpp:672-765, interpol_and_mul_h 778-1070): the same products
 * as or_interpolandmul, every transform in the 4-step form. */
void or_interpolandmul_mt(uint32_t *out, const uint32_t *pA, const uint32_t *pB,
                          const uint32_t *roots, uint32_t log2m, uint32_t nthreads)
{
    const int nthr = nthreads ? (int)nthreads : 1;
    size_t m = (size_t)1 << log2m;
    uint32_t *A = (uint32_t *)malloc(m * 32), *B = (uint32_t *)malloc(m * 32);
    uint32_t *tmp = (uint32_t *)malloc(2 * m * 32);
    memcpy(A, pA, m * 32);
    memcpy(B, pB, m * 32);
#pragma omp parallel for num_threads(nthr) schedule(static)
    for (size_t i = 0; i < m; i++) or_montmult(&out[8 * (2 * i)], &A[8 * i], &B[8 * i], 1);
    uint32_t *X[2] = {A, B};
    for (int t = 0; t < 2; t++) { /* interpol_parallel_odd_h, ntt.cpp:660-670 */
        intt4(X[t], roots, log2m, 2, 1, tmp, nthr);
#pragma omp parallel for num_threads(nthr) schedule(static)
        for (size_t i = 0; i < m; i++) or_montmult(&X[t][8 * i], &X[t][8 * i], &roots[8 * i], 1);
        ntt4(X[t], roots, log2m, 2, 1, tmp, nthr);
    }
#pragma omp parallel for num_threads(nthr) schedule(static)
    for (size_t i = 0; i < m; i++) or_montmult(&out[8 * (2 * i + 1)], &A[8 * i], &B[8 * i], 1);
    intt4(out, roots, log2m + 1, 1, 0, tmp, nthr);
    free(A);
    free(B);
    free(tmp);
}

void or_interpolandmul(uint32_t *out, const uint32_t *pA, const uint32_t *pB,
                       const uint32_t *roots, uint32_t log2m, uint32_t nthreads)
{
    (void)nthreads;
    size_t m = (size_t)1 << log2m;
    uint32_t *A = (uint32_t *)malloc(m * 32), *B = (uint32_t *)malloc(m * 32);
    memcpy(A, pA, m * 32);
    memcpy(B, pB, m * 32);
    for (size_t i = 0; i < m; i++) or_montmult(&out[8 * (2 * i)], &A[8 * i], &B[8 * i], 1);
    interpol_odd(A, roots, log2m);
    interpol_odd(B, roots, log2m);
    for (size_t i = 0; i < m; i++) or_montmult(&out[8 * (2 * i + 1)], &A[8 * i], &B[8 * i], 1);
    or_intt(out, roots, log2m + 1, 1, 0);
    free(A);
    free(B);
}

void or_mpoly_eval_csr(uint32_t *out, const uint64_t *rowptr, const uint32_t *col,
                       const uint32_t *val, const uint32_t *wit, uint32_t nrows)
{
#pragma omp parallel for schedule(static)
    for (uint32_t r = 0; r < nrows; r++) {
        uint32_t acc[8] = {0}, wm[8], t[8];
        for (uint64_t e = rowptr[r]; e < rowptr[r + 1]; e++) {
            or_to_montgomeryN(wm, &wit[8 * (size_t)col[e]], 1, 1);
            or_montmult(t, &val[8 * e], wm, 1);
            or_addm(acc, acc, t, 1);
        }
        memcpy(&out[8 * (size_t)r], acc, 32);
    }
}

/*
 * Closed-form evaluation helpers for the size-independent checks (not a
 * restatement of reference code): they let a test check a 2m-coefficient
 * product polynomial at one point, polH(z) == A(z) B(z), in O(m).
 */

/* Horner value at x (Montgomery) of n normal-form coefficients; normal-form
 * result (a Montgomery product by x R keeps the normal form). */
void or_poly_eval(uint32_t *out, const uint32_t *coeffs, uint64_t n, const uint32_t *x_mont)
{
    uint32_t acc[8] = {0};
    for (uint64_t i = n; i-- > 0;) {
        or_montmult(acc, acc, x_mont, 1);
        or_addm(acc, acc, &coeffs[8 * i], 1);
    }
    memcpy(out, acc, 32);
}

/* Value at z (Montgomery, outside <w_m>) of the degree < m interpolant of the
 * m Montgomery evaluations e_j = A(w^j), by the barycentric form
 *   A(z) = (z^m - 1)/m * sum_j e_j w^j / (z - w^j),
 * the inverses batched (one inversion).  Montgomery result; returns -1 if z
 * lies in the domain or memory runs out. */
int or_lagrange_eval(uint32_t *out, const uint32_t *evals, uint32_t log2m, const uint32_t *z)
{
    size_t m = (size_t)1 << log2m;
    uint32_t *wj = malloc(32 * m), *pre = malloc(32 * m);
    if (!wj || !pre) {
        free(wj);
        free(pre);
        return -1;
    }
    uint32_t w[8], d[8], acc[8], inv[8], t[8];
    memcpy(w, ROOT28_MONT, 32);
    for (uint32_t i = log2m; i < 28; i++) or_montmult(w, w, w, 1);
    memcpy(wj, OR_ONE[1], 32);
    for (size_t j = 1; j < m; j++) or_montmult(&wj[8 * j], &wj[8 * (j - 1)], w, 1);
    memcpy(acc, OR_ONE[1], 32);
    for (size_t j = 0; j < m; j++) {
        or_subm(d, z, &wj[8 * j], 1);
        or_montmult(acc, acc, d, 1);
        memcpy(&pre[8 * j], acc, 32);
    }
    static const uint32_t zero[8] = {0};
    if (memcmp(acc, zero, 32) == 0) {
        free(wj);
        free(pre);
        return -1;
    }
    or_montinv(inv, acc, 1);  /* (prod d_j)^-1 */
    uint32_t sum[8] = {0};
    for (size_t j = m; j-- > 0;) {
        uint32_t invd[8];
        if (j) or_montmult(invd, inv, &pre[8 * (j - 1)], 1);
        else memcpy(invd, inv, 32);
        or_subm(d, z, &wj[8 * j], 1);
        or_montmult(inv, inv, d, 1);
        or_montmult(t, &evals[8 * j], &wj[8 * j], 1);
        or_montmult(t, t, invd, 1);
        or_addm(sum, sum, t, 1);
    }
    /* (z^m - 1) / m */
    uint32_t zm[8], mm[8] = {0}, minv[8];
    memcpy(zm, z, 32);
    for (uint32_t i = 0; i < log2m; i++) or_montmult(zm, zm, zm, 1);
    or_subm(zm, zm, OR_ONE[1], 1);
    mm[log2m >> 5] = 1u << (log2m & 31);
    or_to_montgomeryN(mm, mm, 1, 1);
    or_montinv(minv, mm, 1);
    or_montmult(sum, sum, zm, 1);
    or_montmult(out, sum, minv, 1);
    free(wj);
    free(pre);
    return 0;
}
