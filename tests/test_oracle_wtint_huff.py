"""Pins the oracle's wt_int / wt_huff queries and FM count by brute force.

wt_int (wt_int/wt_int_par.hpp:519-660) and wt_pc (wt_huff/wt_pc_par.hpp:430-554)
answers are fixed by the text itself: rank(i,c) = #c in s[0,i), select(i,c) =
position of the i-th c, access(i) = s[i], lex counts and range counts = counts of
values below / above / inside the bounds.  The Huffman shape
(wt_huff/wt_huff_par.hpp:92-117) is checked against an independent heapq
Huffman cost and the Kraft equality; its BFS node order (sdsl byte_tree) is
parity unpinned.  FM count is checked against naive substring counting and
an independent suffix-array binary search.
"""
import heapq

import numpy as np
import pytest


@pytest.mark.parametrize("n,hi,L", [(1, 2, 1), (500, 3, 2), (3000, 1000, 10), (2000, 300, 12),
                                    (4000, 1 << 16, 16), (777, 5, 5)])
def test_wtint_queries_brute(orc, n, hi, L):
    rng = np.random.default_rng(n + L)
    s = rng.integers(0, hi, n).astype(np.uint32)
    W = orc.WtInt(s, L)
    for i in rng.integers(0, n, 200):
        assert W.access(int(i)) == s[i]
    cs = np.unique(s)
    for _ in range(200):
        c = int(rng.choice(cs)) if rng.random() < 0.9 else int(rng.integers(0, 1 << L))
        i = int(rng.integers(0, n + 1))
        assert W.rank(i, c) == int((s[:i] == c).sum())
        occ = np.nonzero(s == c)[0]
        if occ.size:
            j = int(rng.integers(1, occ.size + 1))
            assert W.select(j, c) == occ[j - 1]
            assert W.select(occ.size + 1, c) == orc.NONE
        else:
            assert W.select(1, c) == orc.NONE
    # lex_count (:712-741), lex_smaller_count (:753-780): counts of values below /
    # above c; range_search_2d count (:801-870): positions in [lb,rb] with values in
    # [vlb,vrb], where the reference clamps vrb to 2^L and, its leaf test not
    # comparing vlb, counts the value 2^L-1 for vlb = 2^L
    top = 1 << L
    for _ in range(200):
        i, j = sorted(int(x) for x in rng.integers(0, n + 1, 2))
        c = int(rng.integers(0, top + 3))
        seg = s[i:j].astype(np.int64)
        inside = c < top
        assert W.lex_count(i, j, c) == (int((s[:i] == c).sum()) if inside else 0, int((seg < c).sum()),
                                        int((seg > c).sum()) if inside else 0)
        assert W.lex_smaller_count(j, c) == (int((s[:j] == c).sum()) if inside else 0, int((s[:j] < c).sum()))
        lb, rb = (int(x) for x in rng.integers(0, n, 2))
        vlb, vrb = sorted(int(x) for x in rng.integers(0, top + 3, 2))
        got = W.range_count(lb, rb, vlb, vrb)
        hi_ = min(vrb, top)
        if lb > rb or vlb > hi_:
            assert got == 0
        else:
            seg = s[lb:rb + 1].astype(np.int64)
            assert got == int(((seg >= min(vlb, top - 1)) & (seg <= hi_)).sum())
    assert W.range_count(0, n - 1, top, top + 7) == int((s == top - 1).sum())
    assert W.range_count(0, n - 1, 3, 2) == 0


def huff_cost(freqs):
    h = [f for f in freqs if f > 0]
    if len(h) <= 1:
        return 0
    heapq.heapify(h)
    cost = 0
    while len(h) > 1:
        a = heapq.heappop(h)
        b = heapq.heappop(h)
        cost += a + b
        heapq.heappush(h, a + b)
    return cost


def dna_text(rng, n, sentinel=True):
    t = rng.choice(np.frombuffer(b"ACGT", np.uint8), n, p=[0.3, 0.2, 0.2, 0.3])
    t[rng.random(n) < 0.01] = ord("N")
    if sentinel:
        t = np.concatenate([t, np.zeros(1, np.uint8)])
    return t


@pytest.mark.parametrize("kind", ["dna", "bytes", "single", "two"])
def test_huff_shape_and_queries(orc, kind):
    rng = np.random.default_rng(hash(kind) & 0xFFFF)
    if kind == "dna":
        t = dna_text(rng, 5000)
    elif kind == "bytes":
        t = rng.zipf(1.5, 6000).clip(0, 255).astype(np.uint8)
    elif kind == "single":
        t = np.full(300, 65, np.uint8)
    else:
        t = rng.choice(np.array([7, 9], np.uint8), 1000)
    H = orc.Huff(t)
    C = np.bincount(t, minlength=256)
    h = H.H
    lens = np.array([h.path_len[c] for c in range(256)])
    present = C > 0
    assert int(H.nbits) == int((C * lens).sum()) == huff_cost(C.tolist())
    if present.sum() >= 2:
        assert abs(sum(2.0 ** -lens[present]) - 1.0) < 1e-12
    for i in rng.integers(0, t.size, 200):
        assert H.access(int(i)) == t[i]
    for _ in range(300):
        c = int(rng.choice(np.nonzero(present)[0]))
        i = int(rng.integers(0, t.size + 1))
        assert H.rank(i, c) == int((t[:i] == c).sum())
        occ = np.nonzero(t == c)[0]
        j = int(rng.integers(1, occ.size + 1))
        assert H.select(j, c) == occ[j - 1]
    assert H.rank(5, 250 if not present[250] else 251) == 0 or present[251]


def test_sa_bwt_and_fm_count(orc):
    rng = np.random.default_rng(11)
    t = dna_text(rng, 3000)
    sa = orc.sa_brute(t)
    tb = t.tobytes()
    assert [tb[i:] for i in sa] == sorted(tb[i:] for i in range(len(tb)))
    bwt = orc.bwt_from_sa(t, sa)
    # fm-index/par/WT.C:38-41: bwt[i] = t[sa[i]-1], or t[n-1] when sa[i] == 0
    exp = np.array([t[i - 1] if i > 0 else t[-1] for i in sa], np.uint8)
    assert np.array_equal(bwt, exp)
    m = 6
    pats = []
    for k in range(300):
        p = int(rng.integers(0, t.size - m - 1))
        pat = t[p:p + m].copy()
        if k % 2:
            pat[int(rng.integers(0, m))] = ord("ACGT"[int(rng.integers(0, 4))])
        pats.append(pat)
    pats = np.array(pats, np.uint8)
    got = orc.fm_count(bwt, pats)
    naive = [sum(1 for i in range(len(tb) - m + 1) if tb[i:i + m] == p.tobytes()) for p in pats]
    assert got.tolist() == naive
    assert orc.sa_count(t, sa, pats).tolist() == naive
