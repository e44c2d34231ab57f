"""Pins the packed bit arithmetic of the demography kernel's census scan
(CPU, exhaustive or brute force); the GPU parity tests check the kernel that
uses it against the oracle.

- Mosquito death test: a slot's 16-bit word holds its death day mod 4096 in
  bits 4-15.  The kernel subtracts the word from ((t mod 4096) << 4 | 15)
  per 16-bit half and reads bit 15: clear iff (t - death day) mod 4096 <
  2048, i.e. the death day is today or in the 2047 days before.
- E -> I due test: (int16)(t - next) > 0, done as bit 15 clear in both
  d = t - next and d - 1 (mod 2^16), per 16-bit half.
- Owner marks of the first 64 birth quads: each lane marks the first quad
  that starts in its four bitmap words, and a quad's owner is the last mark
  at or before it; the word within the owner is the last whose dead-slot
  prefix is at most the quad's first rank.
"""
import numpy as np


def _sub16(a, b):
    return (a.astype(np.int64) - b.astype(np.int64)) & 0xFFFF


def test_mosquito_death_test_all_days():
    t12 = np.arange(4096, dtype=np.int64)[:, None]
    d = np.arange(4096, dtype=np.int64)[None, :]
    expect = ((t12 - d) % 4096) < 2048           # death day within [t - 2047, t]
    for low in (0, 5, 15):                       # status and strain bits below the day
        r = _sub16((t12 << 4) | 15, (d << 4) | low)
        died = (r >> 15) == 0
        assert np.array_equal(died, expect), low
    # the test the kernel used before: bit 11 of ~(t12 + 4096 - d)
    old = (((~((t12 + 4096) - d)) >> 11) & 1) == 1
    assert np.array_equal(old, expect)


def test_mosquito_death_test_two_halves():
    rng = np.random.default_rng(1)
    w = rng.integers(0, 1 << 32, 200_000, dtype=np.uint64)
    t = rng.integers(0, 1 << 23, 200_000, dtype=np.uint64)
    th = ((((t & 0xFFF) << 4) | 0xF) * 0x10001) & 0xFFFFFFFF
    lo = _sub16(th & 0xFFFF, w & 0xFFFF)
    hi = _sub16(th >> 16, w >> 16)
    packed = (lo | (hi << 16)).astype(np.uint64)
    for half in (0, 1):
        day = (w >> (16 * half + 4)) & 0xFFF
        expect = ((t.astype(np.int64) - day.astype(np.int64)) % 4096) < 2048
        assert np.array_equal(((packed >> (16 * half + 15)) & 1) == 0, expect)


def test_exposed_due_test_all_differences():
    d = np.arange(65536, dtype=np.int64)
    expect = d.astype(np.uint16).view(np.int16) > 0
    d1 = (d - 1) & 0xFFFF
    due = (((d | d1) >> 15) & 1) == 0
    assert np.array_equal(due, expect)


def _select_in_word(x, k):
    pos = 0
    for width in (16, 8, 4, 2, 1):
        c = bin(x & ((1 << width) - 1)).count("1")
        if k >= c:
            k -= c
            x >>= width
            pos += width
    return pos


def _owner_marks_case(dead_words, nwords):
    """Birth slots of the first 64 quads as the kernel finds them."""
    words = [dead_words[w] if w < nwords else 0 for w in range(256)]
    pre, wpre = 0, []
    for w in range(256):
        wpre.append(pre)
        pre += bin(words[w]).count("1")
    total = pre
    own = [0] * 64
    for lane in range(64):
        first, end = wpre[4 * lane], (wpre[4 * lane + 4] if lane < 63 else total)
        qa = (first + 3) >> 2
        if 4 * qa < end and qa < 64:
            own[qa] = lane + 1
    out = {}
    for g in range(64):
        if 4 * g >= total:
            break
        j = max(i for i in range(g + 1) if own[i])
        o = own[j] - 1
        p = wpre[4 * o:4 * o + 4]
        rk = 4 * g
        k = sum(1 for x in p[1:] if x <= rk)
        w = 4 * o + k
        out[g] = 32 * w + _select_in_word(words[w], rk - p[k])
    return out, words


def test_owner_marks_against_brute_force():
    rng = np.random.default_rng(7)
    for trial in range(300):
        nwords = int(rng.integers(1, 257))
        kind = trial % 4
        if kind == 0:       # scattered deaths
            dead = [int(x) for x in (rng.random((nwords, 32)) < 0.1) @ (1 << np.arange(32, dtype=np.uint64))]
        elif kind == 1:     # packed alive slots, an empty tail
            cut = int(rng.integers(0, 32 * nwords))
            dead = [sum(1 << b for b in range(32) if 32 * w + b >= cut) for w in range(nwords)]
        elif kind == 2:     # runs of words with no dead slot
            dead = [int(rng.integers(0, 1 << 32)) if rng.random() < 0.2 else 0 for _ in range(nwords)]
        else:               # every slot dead
            dead = [0xFFFFFFFF] * nwords
        got, words = _owner_marks_case(dead, nwords)
        ranks = [32 * w + b for w in range(256) for b in range(32) if (words[w] >> b) & 1]
        for g, slot in got.items():
            assert slot == ranks[4 * g], (trial, g)
        assert len(got) == min(64, (len(ranks) + 3) // 4)
