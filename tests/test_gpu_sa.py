"""Suffix array on the GPU (prefix doubling, wtg_suffix_array_gpu) against the
host SA-IS (wtg_suffix_array, itself pinned by a comparison sort in the CPU
tests) and a direct sort of the suffixes on small texts: identical arrays, and
the BWT built from it (fm-index/par/WT.C:38-41)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def texts(wsfm):
    rng = np.random.default_rng(17)
    yield "dna", wsfm.gen_dna(5, 1 << 20)
    yield "random256", rng.integers(0, 256, 300001).astype(np.uint8)
    yield "zeros_inside", rng.integers(0, 3, 100003).astype(np.uint8)   # real 0 bytes, no sentinel
    yield "all_a", np.full(70000, 97, np.uint8)
    yield "periodic", np.tile(np.frombuffer(b"abcab", np.uint8), 20001)
    yield "repeats", np.concatenate([wsfm.gen_dna(6, 40000)[:-1]] * 3 + [np.zeros(1, np.uint8)])
    # 255 values (8-bit codes, 7 bytes in round 0); 16 values spread over the
    # byte range (5-bit codes, 12 bytes) with long copied stretches
    yield "alpha255", rng.integers(1, 256, 200003).astype(np.uint8)
    a16 = rng.choice(np.arange(0, 256, 17, dtype=np.uint8)[:16], 60000)
    yield "alpha16_repeats", np.concatenate([a16, a16[5000:45000], a16[:20000]])
    for n in (1, 2, 7, 8, 64, 4095, 4096, 4097):
        yield "tiny%d" % n, rng.integers(0, 4, n).astype(np.uint8) + 65


def brute(t):
    s = t.tobytes()
    return np.array(sorted(range(len(s)), key=lambda i: s[i:]), np.int64)


@pytest.mark.parametrize("case", range(16))
def test_gpu_suffix_array(wsfm, orc, ctx, case):
    cases = list(texts(wsfm))
    assert len(cases) == 16
    name, t = cases[case]
    print("sa case", name, t.size, flush=True)
    d = ctx.to_device(t)
    dsa = wsfm.suffix_array_gpu(ctx, d, t.size)
    try:
        got = dsa.get()
        exp = wsfm.suffix_array(t)
        assert np.array_equal(got, exp), name
        if t.size <= 5000:
            assert np.array_equal(got, brute(t)), name
        dbwt = wsfm.bwt_from_sa_gpu(ctx, d, dsa, t.size)
        try:
            assert np.array_equal(dbwt.get(), orc.bwt_from_sa(t, exp)), name
        finally:
            dbwt.free()
    finally:
        dsa.free()
        d.free()


def test_gpu_suffix_array_past_2e24(wsfm, ctx):
    """n = 2^24 + 4097: more than 4096 tiles of 4096 suffixes, so the tile
    prefix carries across its 4096-tile chunks.  DNA with copied stretches."""
    rng = np.random.default_rng(24)
    base = wsfm.gen_dna(8, (1 << 24) + 4097)
    t = base.copy()
    for _ in range(200):  # copied stretches of up to 10^4 bytes
        L = int(rng.integers(100, 10000))
        src, dst = (int(x) for x in rng.integers(0, t.size - L, 2))
        t[dst:dst + L] = base[src:src + L]
    print("sa case past2e24", t.size, flush=True)
    d = ctx.to_device(t)
    dsa = wsfm.suffix_array_gpu(ctx, d, t.size)
    try:
        assert np.array_equal(dsa.get(), wsfm.suffix_array(t))
    finally:
        dsa.free()
        d.free()


def test_bwt_rejects_a_non_suffix_array(wsfm, ctx):
    """An SA entry outside [0, n) is reported with its own message, and the
    flag it raised does not fail the next call."""
    t = wsfm.gen_dna(9, 5000)
    d = ctx.to_device(t)
    sa = wsfm.suffix_array(t)
    bad = sa.copy()
    bad[123] = t.size + 5
    dbad = ctx.to_device(bad)
    try:
        with pytest.raises(wsfm.WtgError, match="not a suffix array"):
            wsfm.bwt_from_sa_gpu(ctx, d, dbad, t.size)
        dsa = wsfm.suffix_array_gpu(ctx, d, t.size)   # a clean call afterwards
        try:
            assert np.array_equal(dsa.get(), sa)
        finally:
            dsa.free()
    finally:
        dbad.free()
        d.free()
