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
    finally:
        wt.free(); ds.free()
