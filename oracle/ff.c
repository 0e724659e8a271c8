/*
 * Oracle field arithmetic -- TEST INFRASTRUCTURE ONLY (see csk_oracle.h).
 *
 * Restates montmult_h (cusnarks src/cuda/ff.cpp:225-388: FIOS Montgomery
 * product, fully reduced output), addm_h / subm_h (ff.cpp:679-744),
 * montinv_h (ff.cpp:745-780), to/from_montgomeryN_h (ff.cpp:496-536,
 * 652-668) and the Fq2 product montmult_ext_h (ff.cpp:389-422, u^2 = -1).
 * The arithmetic is done on 4 x 64-bit limbs (the reference's own non-asm
 * path also uses 64-bit words); the memory layout is the reference's
 * 8 x u32 little-endian.
 */
#include <string.h>
#include "csk_oracle.h"
#include "oracle_internal.h"

typedef unsigned __int128 u128;

const uint64_t OR_P[2][4] = {
    {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
    {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL}};
static const uint64_t OR_NP[2] = {0x87d20782e4866389ULL, 0xc2e1f593efffffffULL};
/* R^2 mod p (constants.cpp:55-58) */
static const uint32_t OR_R2[2][8] = {
    {1401617033, 4079811675, 3561292283, 3051821329, 172064758, 1202396927, 3401069855, 114859889},
    {2921426343, 465102405, 3814480355, 1409170097, 1404797061, 2353627965, 2135835813, 35049649}};
/* R mod p = Montgomery one (constants.cpp:136-139) */
const uint32_t OR_ONE[2][8] = {
    {3314486685, 3546104717, 4123462461, 175696680, 2021213740, 1718526831, 2584207151, 235567041},
    {1342177275, 2895524892, 2673921321, 922515093, 2021213742, 1718526831, 2584207151, 235567041}};

static inline int geq_p(const uint64_t *t, uint32_t pidx)
{
    for (int i = 3; i >= 0; i--) {
        if (t[i] > OR_P[pidx][i]) return 1;
        if (t[i] < OR_P[pidx][i]) return 0;
    }
    return 1;
}

static inline void sub_p(uint64_t *t, uint32_t pidx)
{
    u128 br = 0;
    for (int i = 0; i < 4; i++) {
        u128 d = (u128)t[i] - OR_P[pidx][i] - (uint64_t)br;
        t[i] = (uint64_t)d;
        br = (d >> 64) & 1;
    }
}

void or_montmult(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t pidx)
{
    uint64_t a[4], b[4], t[6] = {0, 0, 0, 0, 0, 0};
    const uint64_t *p = OR_P[pidx];
    memcpy(a, x, 32);
    memcpy(b, y, 32);
    for (int i = 0; i < 4; i++) {
        u128 c = 0;
        for (int j = 0; j < 4; j++) {
            c = (u128)a[j] * b[i] + t[j] + (uint64_t)(c >> 64);
            t[j] = (uint64_t)c;
        }
        c = (u128)t[4] + (uint64_t)(c >> 64);
        t[4] = (uint64_t)c;
        t[5] = (uint64_t)(c >> 64);
        uint64_t m = t[0] * OR_NP[pidx];
        c = (u128)m * p[0] + t[0];
        for (int j = 1; j < 4; j++) {
            c = (u128)m * p[j] + t[j] + (uint64_t)(c >> 64);
            t[j - 1] = (uint64_t)c;
        }
        c = (u128)t[4] + (uint64_t)(c >> 64);
        t[3] = (uint64_t)c;
        t[4] = t[5] + (uint64_t)(c >> 64);
    }
    if (t[4] || geq_p(t, pidx)) sub_p(t, pidx);
    memcpy(z, t, 32);
}

void or_montsquare(uint32_t *z, const uint32_t *x, uint32_t pidx) { or_montmult(z, x, x, pidx); }

void or_addm(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t pidx)
{
    uint64_t a[4], b[4], t[4];
    memcpy(a, x, 32);
    memcpy(b, y, 32);
    u128 c = 0;
    for (int i = 0; i < 4; i++) {
        c = (u128)a[i] + b[i] + (uint64_t)(c >> 64);
        t[i] = (uint64_t)c;
    }
    if ((uint64_t)(c >> 64) || geq_p(t, pidx)) sub_p(t, pidx);
    memcpy(z, t, 32);
}

void or_subm(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t pidx)
{
    uint64_t a[4], b[4], t[4];
    memcpy(a, x, 32);
    memcpy(b, y, 32);
    u128 br = 0;
    for (int i = 0; i < 4; i++) {
        u128 d = (u128)a[i] - b[i] - (uint64_t)br;
        t[i] = (uint64_t)d;
        br = (d >> 64) & 1;
    }
    if (br) {
        u128 c = 0;
        for (int i = 0; i < 4; i++) {
            c = (u128)t[i] + OR_P[pidx][i] + (uint64_t)(c >> 64);
            t[i] = (uint64_t)c;
        }
    }
    memcpy(z, t, 32);
}

/* x^(p-2) in the Montgomery domain: (xR)^-1 R = x^-1 R.  montinv_h returns the
 * same (unique) value through an almost-Montgomery inverse. */
void or_montinv(uint32_t *z, const uint32_t *x, uint32_t pidx)
{
    uint64_t e[4];
    memcpy(e, OR_P[pidx], 32);
    e[0] -= 2; /* no borrow: p is odd and > 2 */
    uint32_t acc[8], base[8];
    memcpy(acc, OR_ONE[pidx], 32);
    memcpy(base, x, 32);
    for (int i = 255; i >= 0; i--) {
        or_montmult(acc, acc, acc, pidx);
        if ((e[i >> 6] >> (i & 63)) & 1) or_montmult(acc, acc, base, pidx);
    }
    memcpy(z, acc, 32);
}

void or_to_montgomeryN(uint32_t *z, const uint32_t *x, uint32_t n, uint32_t pidx)
{
    for (uint32_t i = 0; i < n; i++) or_montmult(&z[8 * i], &x[8 * i], OR_R2[pidx], pidx);
}

void or_from_montgomeryN(uint32_t *z, const uint32_t *x, uint32_t n, uint32_t pidx)
{
    static const uint32_t one[8] = {1, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = 0; i < n; i++) or_montmult(&z[8 * i], &x[8 * i], one, pidx);
}

void or_montmultN(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t n, uint32_t pidx)
{
    for (uint32_t i = 0; i < n; i++) or_montmult(&z[8 * i], &x[8 * i], &y[8 * i], pidx);
}

/* (x0 + x1 u)(y0 + y1 u), u^2 = -1, Karatsuba as mulmontu256_2 / montmult_ext_h */
void or_montmult_ext(uint32_t *z, const uint32_t *x, const uint32_t *y, uint32_t pidx)
{
    uint32_t t0[8], t1[8], s0[8], s1[8], m[8];
    or_montmult(t0, x, y, pidx);
    or_montmult(t1, x + 8, y + 8, pidx);
    or_addm(s0, x, x + 8, pidx);
    or_addm(s1, y, y + 8, pidx);
    or_montmult(m, s0, s1, pidx);
    or_subm(m, m, t0, pidx);
    or_subm(z + 8, m, t1, pidx);
    or_subm(z, t0, t1, pidx);
}
