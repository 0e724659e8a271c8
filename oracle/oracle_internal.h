/* Shared internals of the CPU oracle -- TEST INFRASTRUCTURE ONLY. */
#ifndef CSK_ORACLE_INTERNAL_H
#define CSK_ORACLE_INTERNAL_H
#include <stdint.h>

extern const uint64_t OR_P[2][4];
extern const uint32_t OR_ONE[2][8];

#define OR_MIN(a, b) ((a) < (b) ? (a) : (b))

static inline int or_eq(const uint32_t *a, const uint32_t *b, int nw)
{
    for (int i = 0; i < nw; i++)
        if (a[i] != b[i]) return 0;
    return 1;
}

static inline int or_iszero(const uint32_t *a, int nw)
{
    for (int i = 0; i < nw; i++)
        if (a[i]) return 0;
    return 1;
}

#endif
