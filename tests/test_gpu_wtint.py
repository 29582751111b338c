"""GPU parity of sdsl::wt_int<> (wt_int/wt_int_par.hpp:368-660): tree bits,
rank9 directory and select content bit-exact against the oracle; rank, access,
inverse_select and select answers equal to brute force on the text."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


CASES = [(1, 1, 0, np.uint32), (10, 7, 0, np.uint32), (5000, 1000, 0, np.uint32),
         (5000, 1000, 14, np.uint32), (65536, 256, 0, np.uint8), (100003, 1 << 20, 20, np.uint32),
         (70000, 3, 0, np.uint16), (40000, 1 << 18, 0, np.uint64)]


@pytest.mark.parametrize("n,hi,max_level,dt", CASES)
def test_wtint_vs_oracle(ctx, orc, n, hi, max_level, dt):
    rng = np.random.default_rng(n + hi)
    s = rng.integers(0, hi, n).astype(dt)
    say("wtint", n, hi, max_level, np.dtype(dt).name)
    ds = ctx.to_device(s)
    wt = ctx.wtint_build(ds, n, max_level)
    try:
        L = max_level or (max(1, int(s.max())).bit_length())
        assert wt.L == L and wt.n == n and wt.sigma == len(np.unique(s))
        O = orc.WtInt(s, L)
        assert np.array_equal(wt.words(), O.words)
        assert np.array_equal(wt.directory(), O.D)
        for sel, osel in ((wt.select1(), O.S1), (wt.select0(), O.S0)):
            c = sel.content()
            assert c["arg_cnt"] == osel.arg_cnt
            if osel.sb:
                assert np.array_equal(c["superblock"], osel.superblock[:osel.sb])
                assert np.array_equal(c["mini"], osel.mini[:osel.sb * 64])
        q = min(n, 3000)
        ii = rng.integers(0, n, q).astype(np.uint64)
        assert np.array_equal(wt.access(ii), s[ii.astype(np.int64)].astype(np.uint64))
        r, c = wt.inverse_select(ii)
        assert np.array_equal(c, s[ii.astype(np.int64)].astype(np.uint64))
        exp_r = [int((s[:i] == s[i]).sum()) for i in ii[:300]]
        assert r[:300].tolist() == exp_r
        cs = np.where(rng.random(q) < 0.9, s[rng.integers(0, n, q)], rng.integers(0, 1 << L, q)).astype(np.uint64)
        ri = rng.integers(0, n + 1, q).astype(np.uint64)
        got = wt.rank(ri, cs)
        exp = np.array([O.rank(int(a), int(b)) for a, b in zip(ri, cs)], np.uint64)
        assert np.array_equal(got, exp)
        for a, b, g in list(zip(ri, cs, got))[:200]:
            assert g == int((s[:int(a)] == b).sum())
        # select: valid ranks and one past the count (-> NONE)
        cnt = np.array([O.rank(n, int(b)) for b in cs], np.int64)
        jj = np.array([int(rng.integers(1, k + 1)) if k else 1 for k in cnt], np.uint64)
        jj[::7] = (cnt[::7] + 1).astype(np.uint64)
        got = wt.select(jj, cs)
        exp = np.array([O.select(int(a), int(b)) for a, b in zip(jj, cs)], np.uint64)
        assert np.array_equal(got, exp)
        for a, b, g in list(zip(jj, cs, got))[:200]:
            occ = np.nonzero(s == b)[0]
            assert g == (occ[int(a) - 1] if int(a) <= occ.size else orc.NONE)
        lex_and_range(wt, O, s, L, rng, orc)
    finally:
        wt.free(); ds.free()


def lex_and_range(wt, O, s, L, rng, orc):
    """lex_count / lex_smaller_count / range_search_2d counts against the oracle's
    restatement of the reference traversals and brute force on the text."""
    n = s.size
    q = max(8, min(2000, 4 * n))
    top = 1 << L
    a = rng.integers(0, n + 1, q)
    b = rng.integers(0, n + 1, q)
    i, j = np.minimum(a, b).astype(np.uint64), np.maximum(a, b).astype(np.uint64)
    i[:3], j[:3] = [0, 0, n], [n, 0, n]
    c = np.where(rng.random(q) < 0.8, s[rng.integers(0, n, q)].astype(np.int64) + rng.integers(-1, 2, q),
                 rng.integers(0, top + 3, q))
    c = np.maximum(c, 0).astype(np.uint64)
    rk, sm, gr = wt.lex_count(i, j, c)
    rk2, sm2 = wt.lex_smaller_count(j, c)
    for k in range(q):
        e = O.lex_count(int(i[k]), int(j[k]), int(c[k]))
        assert (int(rk[k]), int(sm[k]), int(gr[k])) == e, k
        assert (int(rk2[k]), int(sm2[k])) == O.lex_smaller_count(int(j[k]), int(c[k])), k
    for k in range(min(q, 300)):
        seg = s[int(i[k]):int(j[k])].astype(np.int64)
        ck = int(c[k])
        assert int(sm[k]) == int((seg < ck).sum()) and int(sm2[k]) == int((s[:int(j[k])] < ck).sum())
        if ck < top:
            assert int(gr[k]) == int((seg > ck).sum()) and int(rk[k]) == int((s[:int(i[k])] == ck).sum())
    # invalid intervals: i > j or j > n
    bad_i = np.array([5 % (n + 1) + 1, 0], np.uint64)
    bad_j = np.array([5 % (n + 1), n + 1], np.uint64)
    r = wt.lex_count(bad_i, bad_j, np.zeros(2, np.uint64))
    assert all(int(x) == orc.NONE for x in np.concatenate(r))
    # range counts, including lb > rb, vlb > vrb, vrb above 2^L and vlb = 2^L
    lb = rng.integers(0, n, q).astype(np.uint64)
    rb = np.minimum(lb + rng.integers(0, max(1, n // 3), q), n - 1).astype(np.uint64)
    rb[:2] = [lb[0], lb[1] - 1 if lb[1] else 0]
    v1 = rng.integers(0, top + 2, q)
    v2 = rng.integers(0, top + 2, q)
    vlb, vrb = np.minimum(v1, v2).astype(np.uint64), np.maximum(v1, v2).astype(np.uint64)
    vlb[2], vrb[2] = 7, 3
    vlb[3], vrb[3] = top, top + 5
    vlb[4], vrb[4] = 0, top + 9
    got = wt.range_count(lb, rb, vlb, vrb)
    for k in range(q):
        assert int(got[k]) == O.range_count(int(lb[k]), int(rb[k]), int(vlb[k]), int(vrb[k])), k
    for k in range(min(q, 300)):
        lo, hi = int(vlb[k]), min(int(vrb[k]), top)
        if lo > hi or int(lb[k]) > int(rb[k]):
            assert int(got[k]) == 0
            continue
        lo = min(lo, top - 1)   # the reference's leaf test does not compare vlb
        seg = s[int(lb[k]):int(rb[k]) + 1].astype(np.int64)
        assert int(got[k]) == int(((seg >= lo) & (seg <= hi)).sum()), k
    assert int(wt.range_count(np.array([0], np.uint64), np.array([n], np.uint64), np.zeros(1, np.uint64),
                              np.ones(1, np.uint64))[0]) == orc.NONE
