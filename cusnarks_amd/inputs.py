"""Seeded synthetic inputs shared by tests, the bench and the oracle checks.

One generator everywhere: SplitMix64 seeded with the workload seed; a field
element takes 4 consecutive 64-bit draws as little-endian limbs, the top limb
masked to 254 bits, and is rejected if >= the modulus (and, for "nonzero"
draws, if 0).  SplitMix64 is counter based (draw i mixes seed + (i+1)*gamma),
so the stream vectorises with numpy.
"""
from __future__ import annotations

import numpy as np

Q = 21888242871839275222246405745257275088696311157297823662689037894645226208583
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, start: int, count: int, out=None, tmp=None) -> np.ndarray:
    """Draws start .. start+count-1 of the SplitMix64 stream of `seed` (into
    out / with scratch tmp, uint64 arrays of at least count, if given)."""
    z = out[:count] if out is not None else np.empty(count, np.uint64)
    t = tmp[:count] if tmp is not None else np.empty(count, np.uint64)
    with np.errstate(over="ignore"):
        z.view(np.int64)[:] = np.arange(start + 1, start + count + 1, dtype=np.int64)
        z *= _GAMMA
        z += np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        for sh, mul in ((30, _M1), (27, _M2), (31, None)):
            np.right_shift(z, np.uint64(sh), out=t)
            z ^= t
            if mul is not None:
                z *= mul
    return z


def _limbs_of(p: int) -> np.ndarray:
    return np.array([(p >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)], dtype=np.uint64)


def _less_than(c: np.ndarray, p: int) -> np.ndarray:
    """Rows of c (m, 4) u64 little-endian that are < p."""
    pl = _limbs_of(p)
    lt = np.zeros(c.shape[0], dtype=bool)
    eq = np.ones(c.shape[0], dtype=bool)
    for i in (3, 2, 1, 0):
        lt |= eq & (c[:, i] < pl[i])
        eq &= c[:, i] == pl[i]
    return lt


def field_elements(seed: int, n: int, modulus: int = R, nonzero: bool = False) -> np.ndarray:
    """n uniform elements of [0, modulus) (or [1, modulus)) as (n, 8) uint32."""
    out = np.empty((n, 4), dtype=np.uint64)
    got = 0
    pos = 0  # index of the next candidate (each candidate uses 4 draws)
    mask = np.uint64((1 << 62) - 1)  # 254 bits total
    chunk = 1 << 20  # candidates per pass; the buffers are reused across passes
    zbuf = np.empty(4 * min(chunk, int(n * 1.35) + 64), np.uint64)
    tbuf = np.empty_like(zbuf)
    while got < n:
        want = min(int((n - got) * 1.35) + 64, chunk)
        c = splitmix64(seed, 4 * pos, 4 * want, out=zbuf, tmp=tbuf).reshape(want, 4)
        c[:, 3] &= mask
        ok = _less_than(c, modulus)
        if nonzero:
            ok &= (c != 0).any(axis=1)
        acc = c[ok]
        take = min(n - got, acc.shape[0])
        if take < acc.shape[0]:
            # position after the last accepted candidate we use
            last = np.flatnonzero(ok)[take - 1]
            pos += int(last) + 1
        else:
            pos += want
        out[got:got + take] = acc[:take]
        got += take
    return out.view(np.uint32).reshape(n, 8)


def to_ints(a: np.ndarray) -> list[int]:
    """(n, 8) uint32 limbs -> Python ints."""
    a64 = np.ascontiguousarray(a, dtype=np.uint32).view(np.uint64).reshape(-1, 4)
    return [int(r[0]) | (int(r[1]) << 64) | (int(r[2]) << 128) | (int(r[3]) << 192) for r in a64]


def from_ints(vals) -> np.ndarray:
    vals = list(vals)
    out = np.zeros((len(vals), 4), dtype=np.uint64)
    for i, v in enumerate(vals):
        for j in range(4):
            out[i, j] = (v >> (64 * j)) & 0xFFFFFFFFFFFFFFFF
    return out.view(np.uint32).reshape(len(vals), 8)


def log_uniform_sizes(seed: int, count: int, lg_lo: float, lg_hi: float) -> list[int]:
    """count sizes round(2^u), u uniform in [lg_lo, lg_hi] (53-bit uniforms from
    SplitMix64 draws 0 .. count-1 of `seed`)."""
    z = splitmix64(seed, 0, count) >> np.uint64(11)
    u = z.astype(np.float64) * (1.0 / (1 << 53))
    return [int(round(2.0 ** (lg_lo + (lg_hi - lg_lo) * x))) for x in u]
