"""CPU emulation of the launch plans and address arithmetic of
csrc/svh_measure.hip: every (kept index, rest index) pair is visited exactly
once, every address stays inside the state, and every partial-sum slot lies
inside the scratch size the library reports."""
import itertools

import pytest

SCAN_TILE = 2048


def deposit_mask(r, mask):
    a = 0
    while mask and r:
        low = mask & (-mask)
        if r & 1:
            a |= low
        r >>= 1
        mask &= mask - 1
    return a


def masked_add(x, d, mask):
    return ((x | (~mask & (2 ** 64 - 1))) + d) & mask


def deposit_kept(m, kept):
    k = len(kept)
    a = 0
    for i, b in enumerate(kept):
        a |= ((m >> (k - 1 - i)) & 1) << b
    return a


def marg_plan(n, k):
    M = 1 << k
    NR = 1 << (n - k)
    if M <= 4096:
        C = 4096 // M
        cap = NR // 1024 if NR >= 1024 else 1
        C = min(C, cap)
        block = True
    else:
        C = 1 if M >= (1 << 18) else (1 << 18) // M
        C = min(C, NR)
        block = False
    return block, M, C, NR // C


def marg_scratch(n, k):
    block, M, C, chunk = marg_plan(n, k)
    return C * M * 8 if C > 1 else 8


def overlap_blocks(n, k):
    P = 1 << (2 * k)
    NR = 1 << (n - k)
    G = min((1 << 18) // P, 1024)
    cap = NR // 64 if NR >= 64 else 1
    return max(1, min(G, cap))


CASES = [(6, [5, 0]), (7, [0]), (9, [3, 7, 1]), (8, list(range(7, -1, -1))), (10, [2, 9, 4, 6, 0, 1, 8]),
         (12, list(range(12))), (11, []), (14, list(range(12, -1, -1))), (13, [5] + list(range(13))[:5] + list(range(6, 13)))]


@pytest.mark.parametrize("n,kept", CASES)
def test_marginal_plan_covers_each_amplitude_once(n, kept):
    k = len(kept)
    kmask = sum(1 << b for b in kept)
    rest = ((1 << n) - 1) & ~kmask
    block, M, C, chunk = marg_plan(n, k)
    assert C * chunk == 1 << (n - k)
    seen = {}
    slots = set()
    if block:
        for blk in range(M * C):
            m, c = blk % M, blk // M
            base = deposit_kept(m, kept)
            slots.add(c * M + m)
            for t in range(256):
                r0 = c * chunk + t
                if t >= chunk:
                    continue
                x = deposit_mask(r0, rest)
                d = deposit_mask(256, rest)
                r = r0
                while r < (c + 1) * chunk:
                    a = base | x
                    assert a == base | deposit_mask(r, rest)
                    seen[a] = seen.get(a, 0) + 1
                    x = masked_add(x, d, rest)
                    r += 256
    else:
        for g in range(M * C):
            m, c = g % M, g // M
            base = deposit_kept(m, kept)
            slots.add(c * M + m)
            x = deposit_mask(c * chunk, rest)
            for _ in range(chunk):
                a = base | x
                seen[a] = seen.get(a, 0) + 1
                x = masked_add(x, 1, rest)
    assert sorted(seen) == list(range(1 << n)) and set(seen.values()) == {1}
    if C > 1:
        assert max(slots) * 8 < marg_scratch(n, k)


def test_marginal_plan_shapes_at_large_sizes():
    for n in (20, 28, 30, 34):
        for k in range(0, n + 1):
            block, M, C, chunk = marg_plan(n, k)
            assert C >= 1 and chunk >= 1 and C * chunk == 1 << (n - k)
            if block:
                assert M * C <= 4096 * 1 or C == 1
            else:
                assert (M * C + 255) // 256 < 2 ** 31
            assert marg_scratch(n, k) <= (1 << 21)


@pytest.mark.parametrize("n,kept", [(6, [0]), (6, [5, 2]), (8, [1, 6, 3]), (9, [8, 7, 0, 4]),
                                    (10, [1, 2, 3, 5, 9]), (11, [0, 2, 4, 6, 8, 10])])
def test_overlap_plan_covers_each_rest_once_per_pair(n, kept):
    k = len(kept)
    kmask = sum(1 << b for b in kept)
    rest = ((1 << n) - 1) & ~kmask
    M, P = 1 << k, 1 << (2 * k)
    G = overlap_blocks(n, k)
    chunk = (1 << (n - k)) // G
    assert G * chunk == 1 << (n - k)
    T = 1 if P >= 256 else 256 // P
    Q = P // 256 if P >= 256 else 1
    visits = {}
    for g, t in itertools.product(range(G), range(256)):
        sub = 0 if P >= 256 else t // P
        pairs = [t + q * 256 for q in range(Q)] if P >= 256 else [t % P]
        if sub >= chunk:
            continue
        x = deposit_mask(g * chunk + sub, rest)
        d = deposit_mask(T, rest)
        r = sub
        while r < chunk:
            for p in pairs:
                assert p < P
                ai = deposit_kept(p // M, kept) | x
                bj = deposit_kept(p % M, kept) | x
                assert ai < (1 << n) and bj < (1 << n)
                visits[(p, g * chunk + r)] = visits.get((p, g * chunk + r), 0) + 1
            x = masked_add(x, d, rest)
            r += T
    assert len(visits) == P * (1 << (n - k)) and set(visits.values()) == {1}
    assert (G - 1) * P + P - 1 < G * P          # partial slots within the scratch size


def test_scan_plan():
    for k in (0, 1, 5, 11, 12, 20):
        M = 1 << k
        nb = (M + SCAN_TILE - 1) // SCAN_TILE
        per = (nb + 255) // 256
        covered = [b for t in range(256) for b in range(t * per, min(t * per + per, nb))]
        assert covered == list(range(nb))
        idx = [blk * SCAN_TILE + t * 8 + e for blk in range(nb) for t in range(256) for e in range(8)]
        assert sorted(i for i in idx if i < M) == list(range(M))
