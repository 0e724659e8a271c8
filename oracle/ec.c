/*
 * Oracle elliptic-curve code -- TEST INFRASTRUCTURE ONLY (see csk_oracle.h).
 *
 * Restates, over G1 (Fq) and G2 (Fq2), the reference CPU multi-exponentiation:
 *   ec_jacadd_h        cusnarks src/cuda/ec.cpp:131-211 (and ec2_ twin 296-380)
 *   ec_jacaddmixed_h   ec.cpp:216-293 (ec2_ twin 382-460)
 *   ec_jacdouble_h     ec.cpp:462-514 (ec2_ twin 516-565)
 *   ec_inittable_h     ec.cpp:574-635 (2^order subset-sum tables, add_last=1)
 *   ec_jacscmul_opt_h  ec.cpp:699-817 (bit-plane accumulation, pippen_conf=0)
 *   ec_jacdouble_finish_h ec.cpp:2004-2049 (Horner over the 256 bit planes)
 *   ec_jacreduce_server_h / ec_jacreduce_batch_h ec.cpp:1536-1795 (thread
 *     partition: multiples of (order<<5)*order points, last thread takes the rest)
 *   ec_jacaddreduce_finish_h ec.cpp:1175-1264, ec_jac2aff_h ec.cpp:963-1012
 * Infinity is detected as the reference does (ec_iseq_h, ec.h:48-59): X == 0
 * and Y == 1*R (the Z coordinate is not looked at).
 */
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <sys/sysinfo.h>
#include "csk_oracle.h"
#include "oracle_internal.h"

#define MAXW 16 /* words per coordinate, G2 */

/* ------------------------------------------------------------ Fq / Fq2 */
static void fmul(uint32_t *z, const uint32_t *x, const uint32_t *y, int ec2)
{
    if (ec2) or_montmult_ext(z, x, y, 0);
    else or_montmult(z, x, y, 0);
}
static void fadd(uint32_t *z, const uint32_t *x, const uint32_t *y, int ec2)
{
    or_addm(z, x, y, 0);
    if (ec2) or_addm(z + 8, x + 8, y + 8, 0);
}
static void fsub(uint32_t *z, const uint32_t *x, const uint32_t *y, int ec2)
{
    or_subm(z, x, y, 0);
    if (ec2) or_subm(z + 8, x + 8, y + 8, 0);
}
static void finv(uint32_t *z, const uint32_t *x, int ec2)
{
    if (!ec2) {
        or_montinv(z, x, 0);
        return;
    }
    uint32_t n0[8], n1[8], d[8], zero[8] = {0};
    or_montmult(n0, x, x, 0);
    or_montmult(n1, x + 8, x + 8, 0);
    or_addm(d, n0, n1, 0);
    or_montinv(d, d, 0);
    or_montmult(z, x, d, 0);
    or_montmult(n1, x + 8, d, 0);
    or_subm(z + 8, zero, n1, 0);
}
static void fone(uint32_t *z, int ec2)
{
    memcpy(z, OR_ONE[0], 32);
    if (ec2) memset(z + 8, 0, 32);
}

static void set_inf(uint32_t *z, int ec2) /* Jacobian (0, 1, 0) */
{
    int nw = ec2 ? 16 : 8;
    memset(z, 0, 3 * nw * 4);
    fone(z + nw, ec2);
}

static int is_inf(const uint32_t *p, int ec2) /* ec_iseq_h / ec2_iseq_h */
{
    int nw = ec2 ? 16 : 8;
    uint32_t one[16];
    fone(one, ec2);
    return or_iszero(p, nw) && or_eq(p + nw, one, nw);
}

/* ------------------------------------------------------------ point formulas */
void or_ec_jacdouble(uint32_t *z, const uint32_t *x, uint32_t ec2)
{
    int nw = ec2 ? 16 : 8;
    if (is_inf(x, ec2)) {
        memmove(z, x, 3 * nw * 4);
        return;
    }
    uint32_t X1[MAXW], Y1[MAXW], Z1[MAXW], Ysq[MAXW], Ysqsq[MAXW], S[MAXW], M[MAXW];
    uint32_t X2[MAXW], Y2[MAXW], Z2[MAXW];
    memcpy(X1, x, nw * 4);
    memcpy(Y1, x + nw, nw * 4);
    memcpy(Z1, x + 2 * nw, nw * 4);
    fmul(Ysq, Y1, Y1, ec2);
    fmul(Ysqsq, Ysq, Ysq, ec2);
    fmul(S, X1, Ysq, ec2);
    fadd(S, S, S, ec2);
    fadd(S, S, S, ec2);
    fmul(M, X1, X1, ec2);
    fadd(X2, M, M, ec2);
    fadd(M, X2, M, ec2);
    fmul(X2, M, M, ec2);
    fmul(Z2, Y1, Z1, ec2);
    fsub(X2, X2, S, ec2);
    fsub(X2, X2, S, ec2);
    fadd(Y2, Ysqsq, Ysqsq, ec2);
    fadd(Y2, Y2, Y2, ec2);
    fadd(Y2, Y2, Y2, ec2);
    fsub(S, S, X2, ec2);
    fmul(S, M, S, ec2);
    fsub(Y2, S, Y2, ec2);
    fadd(Z2, Z2, Z2, ec2);
    memcpy(z, X2, nw * 4);
    memcpy(z + nw, Y2, nw * 4);
    memcpy(z + 2 * nw, Z2, nw * 4);
}

void or_ec_jacadd(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t ec2)
{
    int nw = ec2 ? 16 : 8;
    if (is_inf(x, ec2)) {
        memmove(z, y, 3 * nw * 4);
        return;
    }
    if (is_inf(y, ec2)) {
        memmove(z, x, 3 * nw * 4);
        return;
    }
    uint32_t X1[MAXW], Y1[MAXW], Z1[MAXW], X2[MAXW], Y2[MAXW], Z2[MAXW];
    uint32_t Z1sq[MAXW], Z1cube[MAXW], Z2sq[MAXW], Z2cube[MAXW];
    uint32_t U1[MAXW], U2[MAXW], S1[MAXW], S2[MAXW], Hsq[MAXW], Hcube[MAXW];
    uint32_t X3[MAXW], Y3[MAXW], Z3[MAXW];
    memcpy(X1, x, nw * 4);
    memcpy(Y1, x + nw, nw * 4);
    memcpy(Z1, x + 2 * nw, nw * 4);
    memcpy(X2, y, nw * 4);
    memcpy(Y2, y + nw, nw * 4);
    memcpy(Z2, y + 2 * nw, nw * 4);
    fmul(Z1sq, Z1, Z1, ec2);
    fmul(Z1cube, Z1sq, Z1, ec2);
    fmul(Z2sq, Z2, Z2, ec2);
    fmul(Z2cube, Z2sq, Z2, ec2);
    fmul(U1, X1, Z2sq, ec2);
    fmul(U2, X2, Z1sq, ec2);
    fmul(S1, Y1, Z2cube, ec2);
    fmul(S2, Y2, Z1cube, ec2);
    if (or_eq(U1, U2, nw)) {
        if (!or_eq(S1, S2, nw)) set_inf(z, ec2);
        else {
            uint32_t xc[3 * MAXW];
            memcpy(xc, x, 3 * nw * 4);
            or_ec_jacdouble(z, xc, ec2);
        }
        return;
    }
    fsub(U2, U2, U1, ec2); /* H */
    fsub(S2, S2, S1, ec2); /* R */
    fmul(Hsq, U2, U2, ec2);
    fmul(Hcube, U2, Hsq, ec2);
    fmul(X3, S2, S2, ec2);
    fsub(X3, X3, Hcube, ec2);
    fmul(U1, U1, Hsq, ec2);
    fsub(X3, X3, U1, ec2);
    fsub(X3, X3, U1, ec2);
    fmul(S1, S1, Hcube, ec2);
    fsub(U1, U1, X3, ec2);
    fmul(Y3, S2, U1, ec2);
    fsub(Y3, Y3, S1, ec2);
    fmul(Z3, Z1, Z2, ec2);
    fmul(Z3, Z3, U2, ec2);
    memcpy(z, X3, nw * 4);
    memcpy(z + nw, Y3, nw * 4);
    memcpy(z + 2 * nw, Z3, nw * 4);
}

void or_ec_jacaddmixed(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t ec2)
{
    int nw = ec2 ? 16 : 8;
    if (is_inf(x, ec2)) {
        memmove(z, y, 3 * nw * 4);
        return;
    }
    if (is_inf(y, ec2)) {
        memmove(z, x, 2 * nw * 4);
        fone(z + 2 * nw, ec2);
        return;
    }
    uint32_t X1[MAXW], Y1[MAXW], X2[MAXW], Y2[MAXW], Z2[MAXW];
    uint32_t Z2sq[MAXW], Z2cube[MAXW], U1[MAXW], S1[MAXW], Hsq[MAXW], Hcube[MAXW];
    uint32_t X3[MAXW], Y3[MAXW], Z3[MAXW];
    memcpy(X1, x, nw * 4);
    memcpy(Y1, x + nw, nw * 4);
    memcpy(X2, y, nw * 4);
    memcpy(Y2, y + nw, nw * 4);
    memcpy(Z2, y + 2 * nw, nw * 4);
    fmul(Z2sq, Z2, Z2, ec2);
    fmul(Z2cube, Z2sq, Z2, ec2);
    fmul(U1, X1, Z2sq, ec2);
    fmul(S1, Y1, Z2cube, ec2);
    if (or_eq(U1, X2, nw)) {
        if (!or_eq(S1, Y2, nw)) set_inf(z, ec2);
        else {
            uint32_t yc[3 * MAXW];
            memcpy(yc, y, 3 * nw * 4);
            or_ec_jacdouble(z, yc, ec2);
        }
        return;
    }
    fsub(X2, X2, U1, ec2); /* H */
    fsub(Y2, Y2, S1, ec2); /* R */
    fmul(Hsq, X2, X2, ec2);
    fmul(Hcube, X2, Hsq, ec2);
    fmul(X3, Y2, Y2, ec2);
    fsub(X3, X3, Hcube, ec2);
    fmul(U1, U1, Hsq, ec2);
    fsub(X3, X3, U1, ec2);
    fsub(X3, X3, U1, ec2);
    fmul(S1, S1, Hcube, ec2);
    fsub(U1, U1, X3, ec2);
    fmul(Y3, Y2, U1, ec2);
    fsub(Y3, Y3, S1, ec2);
    fmul(Z3, Z2, X2, ec2);
    memcpy(z, X3, nw * 4);
    memcpy(z + nw, Y3, nw * 4);
    memcpy(z + 2 * nw, Z3, nw * 4);
}

/* ec_jac2aff_h: Z == 0 -> affine infinity (0, 1) */
void or_ec_jac2aff(uint32_t *y, const uint32_t *x, uint32_t n, uint32_t ec2, uint32_t strip_last)
{
    int nw = ec2 ? 16 : 8;
    int od = strip_last ? 2 : 3;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t *P = &x[(size_t)i * 3 * nw];
        uint32_t *O = &y[(size_t)i * od * nw];
        if (or_iszero(P + 2 * nw, nw)) {
            uint32_t inf[3 * MAXW];
            set_inf(inf, ec2);
            memcpy(O, inf, od * nw * 4);
            continue;
        }
        uint32_t zi[MAXW], zi2[MAXW], zi3[MAXW], X[MAXW], Y[MAXW];
        finv(zi, P + 2 * nw, ec2);
        fmul(zi2, zi, zi, ec2);
        fmul(zi3, zi2, zi, ec2);
        fmul(X, P, zi2, ec2);
        fmul(Y, P + nw, zi3, ec2);
        memcpy(O, X, nw * 4);
        memcpy(O + nw, Y, nw * 4);
        if (!strip_last) fone(O + 2 * nw, ec2);
    }
}

/* ------------------------------------------------------------ Straus server */
#define BATCH_SHIFT 5 /* EC_JACREDUCE_BATCH_SIZE, types.h:57 */
#define NPLANES 256   /* EC_JACREDUCE_TABLE_LEN */

static uint32_t clz256(const uint32_t *s) /* msbu256_h (bigint.cpp:216-225) */
{
    uint32_t cnt = 0;
    for (int i = 7; i >= 0; i--) {
        if (s[i]) return cnt + (uint32_t)__builtin_clz(s[i]);
        cnt += 32;
    }
    return 256; /* zero scalar: contributes no bit-plane work */
}

/* ec_inittable_h with add_last = 1: table[j] = sum of the points whose bit is set in j */
static void init_table(uint32_t *tbl, const uint32_t *pts, uint32_t n, uint32_t order, int ec2)
{
    int nw = ec2 ? 16 : 8;
    uint32_t ntab = (n + order - 1) / order, tsize = 1u << order;
    for (uint32_t i = 0; i < ntab; i++) {
        uint32_t *T = &tbl[(size_t)i * tsize * 3 * nw];
        set_inf(T, ec2);
        uint32_t k = 0, last_pow2 = 1;
        for (uint32_t j = 1; j < tsize; j++) {
            uint32_t *E = &T[(size_t)j * 3 * nw];
            if ((j & (j - 1)) == 0) {
                last_pow2 = j;
                if (i * order + k < n) {
                    memcpy(E, &pts[(size_t)(i * order + k) * 2 * nw], 2 * nw * 4);
                    fone(E + 2 * nw, ec2);
                } else set_inf(E, ec2);
                k++;
            } else {
                or_ec_jacadd(E, &T[(size_t)last_pow2 * 3 * nw], &T[(size_t)(j - last_pow2) * 3 * nw], ec2);
            }
        }
    }
}

typedef struct {
    const uint32_t *scl, *pts;
    uint32_t start, last, order;
    int ec2;
    uint32_t *out; /* Jacobian partial */
} straus_arg_t;

static void *straus_worker(void *vp)
{
    straus_arg_t *a = (straus_arg_t *)vp;
    int nw = a->ec2 ? 16 : 8;
    uint32_t order = a->order, tsize = 1u << order;
    uint32_t bsize = (order << BATCH_SHIFT) * order;
    uint32_t *EP = (uint32_t *)malloc((size_t)NPLANES * 3 * nw * 4);
    uint32_t *tbl = (uint32_t *)malloc((size_t)(bsize / order + 1) * tsize * 3 * nw * 4);
    for (int j = 0; j < NPLANES; j++) set_inf(&EP[(size_t)j * 3 * nw], a->ec2);

    for (uint32_t off = a->start; off < a->last; off += bsize) {
        uint32_t ns = OR_MIN(bsize, a->last - off);
        const uint32_t *scl = &a->scl[(size_t)off * 8];
        const uint32_t *pts = &a->pts[(size_t)off * 2 * nw];
        init_table(tbl, pts, ns, order, a->ec2);
        uint32_t ntab = (ns + order - 1) / order;
        for (uint32_t i = 0; i < ntab; i++) {
            uint32_t gs = OR_MIN(order, ns - order * i);
            uint32_t minlz = 256;
            for (uint32_t j = 0; j < gs; j++) {
                uint32_t lz = clz256(&scl[(size_t)(i * order + j) * 8]);
                if (lz < minlz) minlz = lz;
            }
            if (minlz == 256) continue;
            for (int j = 255 - (int)minlz; j >= 0; j--) {
                uint32_t b = 0;
                for (uint32_t g = 0; g < gs; g++)
                    b |= ((scl[(size_t)(i * order + g) * 8 + (j >> 5)] >> (j & 31)) & 1u) << g;
                if (b) {
                    uint32_t *acc = &EP[(size_t)j * 3 * nw];
                    or_ec_jacadd(acc, &tbl[((size_t)i * tsize + b) * 3 * nw], acc, a->ec2);
                }
            }
        }
    }
    /* ec_jacdouble_finish_h: P = sum_j 2^j EP[j] by Horner */
    uint32_t P[3 * MAXW];
    memcpy(P, &EP[(size_t)(NPLANES - 1) * 3 * nw], 3 * nw * 4);
    for (int i = NPLANES - 1; i > 0; i--) {
        or_ec_jacdouble(P, P, a->ec2);
        or_ec_jacadd(P, P, &EP[(size_t)(i - 1) * 3 * nw], a->ec2);
    }
    memcpy(a->out, P, 3 * nw * 4);
    free(EP);
    free(tbl);
    return NULL;
}

int or_ec_jacreduce(uint32_t *out_aff, const uint32_t *scl, const uint32_t *pts,
                    uint32_t n, uint32_t order, uint32_t ec2, uint32_t max_threads,
                    uint32_t *threads_used)
{
    int nw = ec2 ? 16 : 8;
    uint32_t nprocs = (uint32_t)get_nprocs_conf();
    uint32_t nthr = max_threads ? OR_MIN(max_threads, nprocs) : OR_MIN(32u, nprocs);
    if (order < 2 || order > 12) return -1;
    uint32_t bsize = (order << BATCH_SHIFT) * order;
    uint32_t vpt = n;
    /* ec_jacreduce_server_h partition (ec.cpp:1559-1571) */
    if (n >= nthr * bsize) {
        vpt = n / nthr;
        vpt = ((vpt + bsize - 1) / bsize - 1) * bsize;
    } else nthr = 1;
    if (threads_used) *threads_used = nthr;

    pthread_t *th = (pthread_t *)malloc(nthr * sizeof(pthread_t));
    straus_arg_t *args = (straus_arg_t *)malloc(nthr * sizeof(straus_arg_t));
    uint32_t *parts = (uint32_t *)malloc((size_t)nthr * 3 * nw * 4);
    for (uint32_t t = 0; t < nthr; t++) {
        args[t].scl = scl;
        args[t].pts = pts;
        args[t].start = t * vpt;
        args[t].last = (t == nthr - 1) ? n : (t + 1) * vpt;
        args[t].order = order;
        args[t].ec2 = (int)ec2;
        args[t].out = &parts[(size_t)t * 3 * nw];
        if (nthr == 1) straus_worker(&args[t]);
        else pthread_create(&th[t], NULL, straus_worker, &args[t]);
    }
    if (nthr > 1)
        for (uint32_t t = 0; t < nthr; t++) pthread_join(th[t], NULL);
    /* ec_jacaddreduce_finish_h: sequential sum of the thread partials, to affine */
    uint32_t acc[3 * MAXW];
    memcpy(acc, parts, 3 * nw * 4);
    for (uint32_t t = 1; t < nthr; t++) or_ec_jacadd(acc, acc, &parts[(size_t)t * 3 * nw], ec2);
    or_ec_jac2aff(out_aff, acc, 1, ec2, 1);
    free(th);
    free(args);
    free(parts);
    return 0;
}

/* k*P (double-and-add, Jacobian) -> affine: used only to build test inputs */
void or_ec_scmul_aff(uint32_t *out_aff, const uint32_t *scl, const uint32_t *pt_aff,
                     uint32_t n, uint32_t ec2)
{
    int nw = ec2 ? 16 : 8;
#pragma omp parallel for schedule(dynamic, 64)
    for (uint32_t i = 0; i < n; i++) {
        uint32_t acc[3 * MAXW];
        set_inf(acc, (int)ec2);
        const uint32_t *k = &scl[(size_t)i * 8];
        for (int b = 255; b >= 0; b--) {
            or_ec_jacdouble(acc, acc, ec2);
            if ((k[b >> 5] >> (b & 31)) & 1) or_ec_jacaddmixed(acc, pt_aff, acc, ec2);
        }
        or_ec_jac2aff(&out_aff[(size_t)i * 2 * nw], acc, 1, ec2, 1);
    }
}
