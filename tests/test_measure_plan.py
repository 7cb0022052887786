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
    """kind 0: register bins over the whole state (k <= 4); 1: block per
    (m, chunk); 2: thread per (m, chunk)."""
    M = 1 << k
    N = 1 << n
    NR = 1 << (n - k)
    if k <= 4:
        return 0, M, min((N + 255) // 256, 1024), 0
    if M <= 4096:
        C = 4096 // M
        cap = NR // 1024 if NR >= 1024 else 1
        return 1, M, min(C, cap), NR // min(C, cap)
    C = 1 if M >= (1 << 18) else (1 << 18) // M
    C = min(C, NR)
    return 2, M, C, NR // C


def marg_scratch(n, k):
    kind, M, C, chunk = marg_plan(n, k)
    return C * M * 8 if kind == 0 or C > 1 else 8


def overlap_plan(n, k):
    P = 1 << (2 * k)
    NR = 1 << (n - k)
    if k <= 2:
        return True, min((NR + 255) // 256, 1024), 0
    rbt = 1024 >> k
    G = min((1 << 18) // P, 1024)
    cap = NR // rbt if NR >= rbt else 1
    G = max(1, min(G, cap))
    return False, G, NR // G


CASES = [(6, [5, 0]), (7, [0]), (9, [3, 7, 1]), (8, list(range(7, -1, -1))), (10, [2, 9, 4, 6, 0, 1, 8]),
         (12, list(range(12))), (11, []), (14, list(range(12, -1, -1))), (13, [5] + list(range(13))[:5] + list(range(6, 13)))]


@pytest.mark.parametrize("n,kept", CASES)
def test_marginal_plan_covers_each_amplitude_once(n, kept):
    k = len(kept)
    kmask = sum(1 << b for b in kept)
    rest = ((1 << n) - 1) & ~kmask
    kind, M, C, chunk = marg_plan(n, k)
    seen = {}
    slots = set()
    if kind == 0:
        stride = C * 256
        for blk in range(C):
            for t in range(256):
                a = blk * 256 + t
                while a < 1 << n:
                    b = 0
                    for bit in kept:
                        b = (b << 1) | ((a >> bit) & 1)
                    assert b < M
                    seen[a] = seen.get(a, 0) + 1
                    a += stride
            for m in range(M):
                slots.add(blk * M + m)
    elif kind == 1:
        assert C * chunk == 1 << (n - k)
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
        assert C * chunk == 1 << (n - k)
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
    if kind == 0 or C > 1:
        assert max(slots) * 8 < marg_scratch(n, k)


def test_marginal_plan_shapes_at_large_sizes():
    for n in (20, 28, 30, 34):
        for k in range(0, n + 1):
            kind, M, C, chunk = marg_plan(n, k)
            assert C >= 1
            if kind:
                assert chunk >= 1 and C * chunk == 1 << (n - k)
            if kind == 2:
                assert (M * C + 255) // 256 < 2 ** 31
            assert marg_scratch(n, k) <= (1 << 21)


@pytest.mark.parametrize("n,kept", [(6, [0]), (6, [5, 2]), (8, [1]), (9, [8, 0]), (8, [1, 6, 3]),
                                    (9, [8, 7, 0, 4]), (10, [1, 2, 3, 5, 9]), (11, [0, 2, 4, 6, 8, 10]),
                                    (7, [6, 5, 4, 3, 2, 1])])
def test_overlap_plan_covers_each_rest_once_per_pair(n, kept):
    """Mirror of k_overlap_reg (k <= 2) and k_overlap_lds (k >= 3): every
    (pair, rest index) is accumulated exactly once and every address and LDS
    slot is in range."""
    k = len(kept)
    kmask = sum(1 << b for b in kept)
    rest = ((1 << n) - 1) & ~kmask
    M, P = 1 << k, 1 << (2 * k)
    NR = 1 << (n - k)
    reg, G, chunk = overlap_plan(n, k)
    visits = {}
    if reg:
        stride = G * 256
        for g, t in itertools.product(range(G), range(256)):
            r = g * 256 + t
            if r >= NR:
                continue
            x = deposit_mask(r, rest)
            d = deposit_mask(stride, rest)
            while r < NR:
                for i in range(M):
                    assert (deposit_kept(i, kept) | x) < 1 << n
                for p in range(P):
                    visits[(p, r)] = visits.get((p, r), 0) + 1
                x = masked_add(x, d, rest)
                r += stride
    else:
        assert G * chunk == NR
        RBT = 1024 // M
        LPT = M * RBT // 256
        T = 1 if P >= 256 else 256 // P
        Q = P // 256 if P >= 256 else 1
        nb = (chunk + RBT - 1) // RBT
        for g in range(G):
            xb = deposit_mask(g * chunk, rest)
            for bt in range(nb):
                rows = set()
                for t in range(256):
                    rr = t % RBT
                    live = bt * RBT + rr < chunk
                    x = masked_add(xb, deposit_mask(rr, rest), rest)
                    assert x == deposit_mask(g * chunk + bt * RBT + rr, rest) or not live
                    for l in range(LPT):
                        i = t // RBT + l * (256 // RBT)
                        assert i < M and i * (RBT + 1) + rr < M * (RBT + 1)
                        rows.add((i, rr))
                        if live:
                            assert (deposit_kept(i, kept) | x) < 1 << n
                assert len(rows) == M * RBT             # the batch fills the whole LDS slab
                QI = 4 if Q >= 16 else (2 if Q >= 4 else 1)
                QJ = Q // QI
                for t in range(256):
                    sub = 0 if P >= 256 else t // P
                    if P >= 256:
                        i0, j0 = (t // (M // QJ)) * QI, (t % (M // QJ)) * QJ
                        pairs = [(i0 + x) * M + j0 + y for x in range(QI) for y in range(QJ)]
                    else:
                        pairs = [t % P]
                    for c in range(sub, RBT, T):
                        r = bt * RBT + c
                        if r >= chunk:
                            continue
                        for p in pairs:
                            assert p < P
                            visits[(p, g * chunk + r)] = visits.get((p, g * chunk + r), 0) + 1
                xb = masked_add(xb, deposit_mask(RBT, rest), rest)
    assert len(visits) == P * NR and set(visits.values()) == {1}


def test_scan_plan():
    for k in (0, 1, 5, 11, 12, 20):
        M = 1 << k
        nb = (M + SCAN_TILE - 1) // SCAN_TILE
        per = (nb + 255) // 256
        covered = [b for t in range(256) for b in range(t * per, min(t * per + per, nb))]
        assert covered == list(range(nb))
        idx = [blk * SCAN_TILE + t * 8 + e for blk in range(nb) for t in range(256) for e in range(8)]
        assert sorted(i for i in idx if i < M) == list(range(M))
