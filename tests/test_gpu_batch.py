"""Batched level-wise trees on the GPU: every sequence's levels must equal the
oracle's serialWT restatement of WT(s_i, n_i, a_i) (wavelet/serialWT/WT.C:32-103)
bit for bit, through the hipGraph replay and through plain stream launches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def check_all(wsfm, orc, B, sym, n, sigma, off):
    for i in range(n.size):
        L = orc.levels_of_sigma(int(sigma[i]))
        assert B.seq(i)[0] == L
        s = sym[int(off[i]):int(off[i]) + int(n[i])]
        exp, _ = orc.levels_serial(s, L) if n[i] else (np.zeros(0, np.uint64), None)
        got = B.extract(i)
        assert np.array_equal(got, exp[:got.size]), (i, int(n[i]), int(sigma[i]))


@pytest.mark.parametrize("seed,nseq,lmax", [(1, 40, 12), (2, 300, 15)])
def test_batch_matches_serial(wsfm, orc, ctx, seed, nseq, lmax):
    print("batch case", seed, nseq, lmax, flush=True)
    rng = np.random.default_rng(seed)
    n = np.floor(2.0 ** rng.uniform(0, lmax, nseq)).astype(np.uint64)
    n[::17] = 0
    sigma = rng.choice([1, 2, 3, 4, 7, 20, 200, 256], nseq).astype(np.uint32)
    sym = np.concatenate([rng.integers(0, s, int(k)).astype(np.uint8) for s, k in zip(sigma, n)] +
                         [np.zeros(0, np.uint8)])
    off = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.uint64)
    d = ctx.to_device(sym if sym.size else np.zeros(1, np.uint8))
    B = wsfm.Batch(ctx, d, n, sigma, off)
    try:
        B.run()
        w1 = B.words()
        check_all(wsfm, orc, B, sym, n, sigma, off)
        B.run()                      # graph replay
        assert np.array_equal(B.words(), w1)
        B.set_graph(False)           # plain launches
        B.run()
        assert np.array_equal(B.words(), w1)
    finally:
        B.free()
        d.free()


def test_batch_generator_workload_small(wsfm, orc, ctx):
    # the batched workload's generator (alphabets 4, 20, 256), scaled down
    n, sigma, sym = wsfm.gen_batch(7, 64, 4.0, 14.0)
    d = ctx.to_device(sym)
    B = wsfm.Batch(ctx, d, n, sigma)
    try:
        B.run()
        off = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.uint64)
        check_all(wsfm, orc, B, sym, n, sigma, off)
    finally:
        B.free()
        d.free()


def test_batch_permuted_offsets_and_bad_symbol(wsfm, orc, ctx):
    rng = np.random.default_rng(9)
    n = np.array([100, 5000, 3, 777], np.uint64)
    sigma = np.array([4, 256, 2, 20], np.uint32)
    parts = [rng.integers(0, s, int(k)).astype(np.uint8) for s, k in zip(sigma, n)]
    order = [2, 0, 3, 1]                       # sequences stored out of order
    sym = np.concatenate([parts[i] for i in order])
    off = np.zeros(4, np.uint64)
    pos = 0
    for i in order:
        off[i] = pos
        pos += int(n[i])
    d = ctx.to_device(sym)
    B = wsfm.Batch(ctx, d, n, sigma, off)
    try:
        B.run()
        check_all(wsfm, orc, B, sym, n, sigma, off)
    finally:
        B.free()
    bad = sym.copy()
    bad[int(off[0])] = 9                        # >= 2^L for the alphabet-4 sequence
    d2 = ctx.to_device(bad)
    B = wsfm.Batch(ctx, d2, n, sigma, off)
    try:
        with pytest.raises(wsfm.WtgError):
            B.run()
    finally:
        B.free()
        d.free()
        d2.free()


def test_batch_plans_and_autotuner(wsfm, orc, ctx):
    # the same sequences under other group sizes (32-bit keys in one group per
    # width, smaller 16-bit-key groups) and under the autotuner's plan: every
    # sequence's levels are unchanged
    rng = np.random.default_rng(4)
    nseq = 120
    n = np.floor(2.0 ** rng.uniform(4, 13, nseq)).astype(np.uint64)
    sigma = rng.choice([4, 20, 256, 3], nseq).astype(np.uint32)
    sym = np.concatenate([rng.integers(0, s, int(k)).astype(np.uint8) for s, k in zip(sigma, n)])
    off = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.uint64)
    d = ctx.to_device(sym)
    ref = wsfm.Batch(ctx, d, n, sigma, off)
    ref.run()
    want = [ref.extract(i) for i in range(nseq)]
    ref.free()
    wide = np.full(9, 255, np.uint8)
    small = np.full(9, 255, np.uint8)
    byte = np.full(9, 255, np.uint8)
    for L in range(1, 9):
        wide[L] = 24 - L
        small[L] = max(0, 14 - L)
        byte[L] = max(2, 8 - L) if L <= 6 else 255   # 8-bit keys where at most 64 groups
    plan, ms = wsfm.Batch.tune(ctx, d, n, sigma, off, reps=1)
    widths = sorted({orc.levels_of_sigma(int(s)) for s in sigma})
    for L in widths:
        assert plan[L] != 255 and (ms[L] >= 0).any()   # tried and chosen
    for p in (wide, small, byte, plan):
        B = wsfm.Batch(ctx, d, n, sigma, off, plan=p)
        try:
            B.run()
            for i in range(nseq):
                assert np.array_equal(B.extract(i), want[i]), (list(p), i)
        finally:
            B.free()
    d.free()
