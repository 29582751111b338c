"""GPU parity of (b) the rank_support_v<1,1> directory
(rank/rank_support_v_par.hpp:69-142) and (b') select_support_mcl<b,1>
(select/select_support_mcl_par.hpp:227-611) against the oracle: the directory
words and the select content (superblocks, mini offsets, long blocks, widths)
bit-exact, and every query answer equal to brute force."""
import numpy as np
import pytest

from test_oracle_rank_select import bits_of, rand_words

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def check_select(sel, osel):
    c = sel.content()
    assert c["arg_cnt"] == osel.arg_cnt and c["sb"] == osel.sb
    assert c["logn"] == osel.logn and c["logn4"] == osel.logn4
    assert c["n_long"] == osel.n_long
    sb = osel.sb
    if sb == 0:
        return
    assert np.array_equal(c["superblock"], osel.superblock[:sb])
    assert np.array_equal(c["mini"], osel.mini[:sb * 64])
    assert np.array_equal(c["mini_width"], osel.mini_width[:sb])
    assert np.array_equal(c["long_width"], osel.long_width[:sb])
    assert np.array_equal(c["long_index"], osel.long_index[:sb])
    if osel.n_long:
        assert np.array_equal(c["long_pool"], osel.long_pool[:osel.n_long * 4096])


CASES = [(1, 0.5, False), (64, 0.5, False), (100, 0.5, False), (511, 0.9, False),
         (512, 0.5, False), (513, 0.1, False), (4096 * 8, 0.5, False), (70001, 0.05, False),
         (1 << 20, 0.01, False), (1 << 20, 0.3, True), (300000, 0.002, False),
         ((1 << 22) + 77, 0.5, False), (1 << 21, 0.999, False), (200000, 0.0, False),
         (200000, 1.0, False)]


@pytest.mark.parametrize("nbits,density,runs", CASES)
def test_rank9_select_build_vs_oracle(ctx, orc, nbits, density, runs):
    rng = np.random.default_rng(nbits + int(density * 1000))
    w = rand_words(rng, nbits, density, runs)
    say("rank_select", nbits, density, runs)
    dw = ctx.to_device(w)
    D, s1, s0 = ctx.rank9_select_build(dw, nbits)
    try:
        assert np.array_equal(D.get(), orc.rank9_build(w, nbits))
        check_select(s1, orc.Select(w, nbits, 1))
        check_select(s0, orc.Select(w, nbits, 0))
        # rank queries: every index for small vectors, random sample otherwise
        pref = np.concatenate([[0], np.cumsum(bits_of(w, nbits))]).astype(np.uint64)
        idx = np.arange(nbits + 1, dtype=np.uint64) if nbits < 5000 else \
            np.unique(np.concatenate([[0, nbits], rng.integers(0, nbits + 1, 20000)])).astype(np.uint64)
        di = ctx.to_device(idx)
        r = ctx.rank9_batch(dw, D, nbits, di)
        assert np.array_equal(r.get(), pref[idx.astype(np.int64)])
        di.free(); r.free()
        for sel, bit in ((s1, 1), (s0, 0)):
            pos = np.nonzero(bits_of(w, nbits) == bit)[0].astype(np.uint64)
            if pos.size == 0:
                continue
            q = np.unique(np.concatenate([[1, pos.size], rng.integers(1, pos.size + 1, 20000)])).astype(np.uint64)
            dq = ctx.to_device(q)
            out = ctx.select_batch(sel, dw, dq)
            assert np.array_equal(out.get(), pos[(q - 1).astype(np.int64)])
            dq.free(); out.free()
    finally:
        D.free(); s1.free(); s0.free(); dw.free()


@pytest.mark.parametrize("bit", [0, 1])
def test_select_from_directory_matches_fused(ctx, orc, bit):
    rng = np.random.default_rng(77 + bit)
    nbits = 3_000_000
    w = rand_words(rng, nbits, 0.02 if bit else 0.98)
    dw = ctx.to_device(w)
    D = ctx.rank9_build(dw, nbits)
    sel = ctx.select_build(dw, D, nbits, bit)
    try:
        check_select(sel, orc.Select(w, nbits, bit))
    finally:
        sel.free(); D.free(); dw.free()


def test_rank_select_large_properties(ctx, wsfm):
    """2^31 bits: rank(select(j)) == j-1 and bit(select(j)) == 1 on samples,
    directory totals equal popcount of the words."""
    nbits = 1 << 31
    w = wsfm.gen_uniform(9, nbits // 64, 1 << 63, np.uint64)
    w ^= wsfm.gen_uniform(10, nbits // 64, 1 << 63, np.uint64) << np.uint64(1)
    say("large", nbits)
    dw = ctx.to_device(w)
    D, s1, s0 = ctx.rank9_select_build(dw, nbits)
    try:
        ones = int(np.unpackbits(w.view(np.uint8)).sum())
        assert s1.arg_cnt == ones and s0.arg_cnt == nbits - ones
        rng = np.random.default_rng(0)
        j = rng.integers(1, ones + 1, 100000).astype(np.uint64)
        dj = ctx.to_device(j)
        p = ctx.select_batch(s1, dw, dj)
        r = ctx.rank9_batch(dw, D, nbits, p)
        pv = p.get()
        assert np.array_equal(r.get(), j - 1)
        assert (((w[(pv >> 6).astype(np.int64)] >> (pv & 63)) & 1) == 1).all()
        for a in (dj, p, r):
            a.free()
    finally:
        D.free(); s1.free(); s0.free(); dw.free()


@pytest.mark.parametrize("nbits", [1, 63, 64, 65, 4099, 100003, (1 << 22) + 77])
@pytest.mark.parametrize("pattern", [10, 1])
def test_pattern_words_vs_oracle(ctx, orc, nbits, pattern):
    """wtg_pattern_words (two-bit patterns) = the oracle's map10 / map01 words,
    both clip modes, with the last bit set (a `10` in the padding)."""
    from test_oracle_rank_select import pattern_case
    say("pattern_words", nbits, pattern)
    w = pattern_case(nbits, nbits * 7 + pattern, True)
    dw = ctx.to_device(w)
    for clip in (False, True):
        got = ctx.pattern_words(dw, nbits, pattern, clip).get()
        assert np.array_equal(got, orc.pattern_words(w, nbits, pattern, clip)), clip
