"""wt_huff on the GPU (wtg_huff_build and its queries) and FM-index count.

Shape, tree bits and rank directory must equal the oracle's restatement of
wt_huff/wt_huff_par.hpp:92-117 + wt_pc_par.hpp:193-257 bit for bit; query
answers must equal brute force on the text; FM counts must equal both the
oracle's backward search and a binary search over the suffix array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fib_text(rng, k):
    # counts 1,1,2,3,5,... -> a Huffman tree of depth k-1
    cnt = [1, 1]
    while len(cnt) < k:
        cnt.append(cnt[-1] + cnt[-2])
    t = np.concatenate([np.full(c, 10 + j, np.uint8) for j, c in enumerate(cnt)])
    return rng.permutation(t)


def texts(wsfm):
    rng = np.random.default_rng(7)
    yield "dna", wsfm.gen_dna(3, 1 << 18)
    yield "zipf", wsfm.gen_zipf_bytes(4, 150001)
    yield "uniform256", rng.integers(0, 256, 70001).astype(np.uint8)
    yield "fib24", fib_text(rng, 24)
    yield "single", np.full(5000, 77, np.uint8)
    yield "two", (rng.random(3001) < 0.01).astype(np.uint8) + 5
    for n in (1, 2, 63, 64, 65, 1023, 1024, 1025, 4097):
        yield "tiny%d" % n, rng.integers(0, 4, n).astype(np.uint8) * 3 + 60
    # alphabet sizes around the count kernel's per-lane counter widths
    # (8, 16 and 32 inner nodes) and the emit's spread counts (up to 6 inner
    # nodes), skewed so the trees are unbalanced
    for sig, n in ((9, 20011), (10, 3000), (17, 50007), (33, 65536 + 129), (34, 30001),
                   (5, 40009), (6, 4099), (7, 70001), (8, 9001)):
        p = 1.0 / np.arange(1, sig + 1) ** 0.7
        yield "alpha%d" % sig, (rng.choice(sig, n, p=p / p.sum()) * 7 + 1).astype(np.uint8)


NTEXTS = 24


def check_shape(wt, oh):
    S, H = wt.shape, oh.H
    assert S.sigma == H.sigma and S.n_nodes == H.n_nodes and S.bv_size == H.bv_size
    nn = H.n_nodes
    assert list(S.bv_pos[:nn]) == list(H.bv_pos[:nn])
    assert [S.child[v][0] for v in range(nn)] == list(H.child0[:nn])
    assert [S.child[v][1] for v in range(nn)] == list(H.child1[:nn])
    assert list(S.parent[:nn]) == list(H.parent[:nn])
    assert list(S.sym[:nn]) == list(H.sym[:nn])
    assert list(S.c_to_leaf) == list(H.c_to_leaf)
    assert list(S.path) == list(H.path) and list(S.path_len) == list(H.path_len)
    cnt = np.bincount(oh.t, minlength=256)
    assert list(S.C) == [0] + list(np.cumsum(cnt))


def brute(t, rng, nq):
    n = t.size
    syms = np.unique(t)
    i = rng.integers(0, n + 1, nq).astype(np.uint64)
    i[:4] = [0, n, n // 2, max(n - 1, 0)]
    c = np.where(rng.random(nq) < 0.9, rng.choice(syms, nq), rng.integers(0, 256, nq)).astype(np.uint32)
    pref = {int(s): np.concatenate([[0], np.cumsum(t == s)]) for s in syms}
    rank = np.array([pref[int(cc)][int(ii)] if int(cc) in pref else 0 for ii, cc in zip(i, c)],
                    np.uint64)
    return i, c, rank


@pytest.mark.parametrize("case", range(NTEXTS))
def test_huff_build_and_queries(wsfm, orc, ctx, case):
    cases = list(texts(wsfm))
    assert len(cases) == NTEXTS
    name, t = cases[case]
    print("huff case", name, t.size, flush=True)
    rng = np.random.default_rng(case)
    d = ctx.to_device(t)
    wt = ctx.huff_build(d)
    try:
        oh = orc.Huff(t)
        check_shape(wt, oh)
        nb = oh.nbits
        words = wt.words()
        assert np.array_equal(words[:(nb + 63) // 64], oh.words)
        assert words[-1] == 0
        if nb:
            assert np.array_equal(wt.directory(), oh.D)
            ni = [v for v in range(oh.H.n_nodes) if oh.H.child0[v] >= 0]
            pref = np.concatenate([[0], np.cumsum(np.unpackbits(oh.words.view(np.uint8),
                                                                 bitorder="little")[:nb])])
            assert [wt.shape.bv_pos_rank[v] for v in ni] == [int(pref[oh.H.bv_pos[v]]) for v in ni]
        # rank
        nq = 4000
        i, c, rank = brute(t, rng, nq)
        assert np.array_equal(wt.rank(i, c), rank)
        assert wt.rank(np.array([t.size + 1], np.uint64), np.array([t[0]], np.uint32))[0] == orc.NONE
        # access / inverse_select
        ia = rng.integers(0, t.size, nq).astype(np.uint64)
        assert np.array_equal(wt.access(ia), t[ia].astype(np.uint64))
        rk, sym = wt.inverse_select(ia)
        assert np.array_equal(sym, t[ia].astype(np.uint64))
        exp = np.array([int((t[:int(k)] == t[int(k)]).sum()) for k in ia[:500]], np.uint64)
        assert np.array_equal(rk[:500], exp)
        assert wt.access(np.array([t.size], np.uint64))[0] == orc.NONE
        # select: i-th occurrence, including out of range
        syms, counts = np.unique(t, return_counts=True)
        cs = rng.choice(syms, nq).astype(np.uint32)
        cnt = dict(zip(syms.tolist(), counts.tolist()))
        js = np.array([rng.integers(0, cnt[int(x)] + 2) for x in cs], np.uint64)
        got = wt.select(js, cs)
        occ = {int(s): np.nonzero(t == s)[0] for s in syms}
        for j, x, g in zip(js[:1500], cs[:1500], got[:1500]):
            j, x = int(j), int(x)
            e = occ[x][j - 1] if 1 <= j <= cnt[x] else orc.NONE
            assert int(g) == int(e), (j, x, g, e)
        # oracle's own query restatement agrees on a few
        for k in range(50):
            assert oh.rank(int(i[k]), int(c[k])) == rank[k]
            assert oh.access(int(ia[k])) == t[int(ia[k])]
    finally:
        wt.free()
        d.free()


def test_huff_unaligned_text(wsfm, orc, ctx):
    t = wsfm.gen_dna(9, 100003)
    big = np.concatenate([np.zeros(3, np.uint8), t])
    d = ctx.to_device(big)
    h = C.c_void_p()
    wsfm._check(wsfm.lib().wtg_huff_build(ctx.h, C.c_void_p(d.ptr.value + 3), t.size, C.byref(h)))
    wt = wsfm.WtHuff(ctx, h)
    try:
        oh = orc.Huff(t)
        assert np.array_equal(wt.words()[:(oh.nbits + 63) // 64], oh.words)
    finally:
        wt.free()
        d.free()


def test_huff_empty(wsfm, ctx):
    d = ctx.empty(1, np.uint8)
    wt = ctx.huff_build(d, 0)
    try:
        assert wt.shape.sigma == 0 and wt.shape.bv_size == 0
        with pytest.raises(wsfm.WtgError):
            wt.rank(np.array([0], np.uint64), np.array([0], np.uint32))
    finally:
        wt.free()
        d.free()


@pytest.mark.parametrize("n,m", [(1 << 16, 8), (1 << 18, 20)])
def test_fm_count(wsfm, orc, ctx, n, m):
    print("fm case", n, m, flush=True)
    t = wsfm.gen_dna(11, n)
    sa = wsfm.suffix_array(t)
    bwt = wsfm.bwt_from_sa(t, sa)
    assert np.array_equal(bwt, orc.bwt_from_sa(t, sa))
    pats = wsfm.gen_patterns(12, t, m, 20000)
    extra = np.array([list(b"ACGT" * 8)[:m], [ord("N")] * m, [0] * m], np.uint8)
    pats = np.concatenate([pats, extra])
    d = ctx.to_device(bwt)
    wt = ctx.huff_build(d)
    try:
        got = wt.fm_count(pats)
        assert np.array_equal(got, orc.sa_count(t, sa, pats))
        assert np.array_equal(got, orc.fm_count(bwt, pats))
        assert (got[0:20000:2] >= 1).all()   # exact substrings occur
    finally:
        wt.free()
        d.free()


def test_fm_count_single_char_patterns(wsfm, orc, ctx):
    t = wsfm.gen_dna(13, 5000)
    sa = wsfm.suffix_array(t)
    bwt = wsfm.bwt_from_sa(t, sa)
    pats = np.arange(256, dtype=np.uint8).reshape(256, 1)
    d = ctx.to_device(bwt)
    wt = ctx.huff_build(d)
    try:
        got = wt.fm_count(pats)
        assert np.array_equal(got, np.bincount(t, minlength=256).astype(np.uint64))
    finally:
        wt.free()
        d.free()


def test_huff_past_2e25_bytes(wsfm, orc, ctx):
    """A text over 2^25 bytes: the histogram's waves then stride over several
    8 KiB chunks each (1024-block grid), so the multi-chunk accumulation runs.
    Skewed bytes (one dominant value) plus a sparse tail."""
    rng = np.random.default_rng(25)
    n = (1 << 25) + 4099
    p = 1.0 / np.arange(1, 41) ** 1.6
    t = (rng.choice(40, n, p=p / p.sum()) * 5 + 3).astype(np.uint8)
    print("huff case past2e25", n, flush=True)
    d = ctx.to_device(t)
    wt = ctx.huff_build(d)
    try:
        cnt = np.bincount(t, minlength=256)
        assert list(wt.shape.C) == [0] + list(np.cumsum(cnt))
        oh = orc.Huff(t)
        check_shape(wt, oh)
        assert np.array_equal(wt.words()[:(oh.nbits + 63) // 64], oh.words)
        syms = [int(t[0]), int(t[n // 2]), int(np.argmax(cnt)), int(np.nonzero(cnt)[0][-1])]
        for s in syms:
            pref = np.concatenate([[0], np.cumsum(t == s, dtype=np.uint64)])
            i = rng.integers(0, n + 1, 2000).astype(np.uint64)
            i[:3] = [0, n, (1 << 25) + 17]
            got = wt.rank(i, np.full(i.size, s, np.uint32))
            assert np.array_equal(got, pref[i.astype(np.int64)]), s
        ia = rng.integers(0, n, 4000).astype(np.uint64)
        assert np.array_equal(wt.access(ia), t[ia].astype(np.uint64))
    finally:
        wt.free()
        d.free()
