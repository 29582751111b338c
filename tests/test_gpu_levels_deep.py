"""Level-wise trees of 25..64 levels (wt_int over wide values: the reference's
wt_int takes any max_level up to 64, wt_int/wt_int_par.hpp:384-395): the bits
equal the oracle's serialWT restatement (up to 63 levels, sparse node chunks)
and an independent numpy closed form (every level = the stable sort by the
top l bits), and the wt_int answers equal brute force on the text."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def closed_form(s, L):
    """Level l = bit L-1-l of the keys stably sorted by their top l bits."""
    s = s.astype(np.uint64)
    n = s.size
    bits = np.zeros(n * L, np.uint8)
    for l in range(L):
        key = np.zeros(n, np.uint64) if l == 0 else (s >> np.uint64(L - l))
        order = np.argsort(key, kind="stable")
        bits[l * n:(l + 1) * n] = ((s[order] >> np.uint64(L - 1 - l)) & np.uint64(1)).astype(np.uint8)
    nb = n * L
    out = np.zeros((nb + 63) // 64 * 8, np.uint8)
    out[:(nb + 7) // 8] = np.packbits(bits, bitorder="little")
    return out.view(np.uint64)


def cases():
    rng = np.random.default_rng(2564)
    yield "u32_L25", rng.integers(0, 1 << 25, 100003, dtype=np.uint64).astype(np.uint32), 25
    yield "u32_L32", rng.integers(0, 1 << 32, 70001, dtype=np.uint64).astype(np.uint32), 32
    yield "u64_L40", rng.integers(0, 1 << 40, 50001, dtype=np.uint64), 40
    yield "u64_L63", rng.integers(0, 1 << 63, 30001, dtype=np.uint64), 63
    yield "u64_L64", rng.integers(0, np.iinfo(np.uint64).max, 20001, dtype=np.uint64), 64
    g = np.minimum(rng.geometric(1e-3, 60007), (1 << 30) - 1).astype(np.uint32)
    yield "geometric_L30", g, 30
    yield "equal_L33", np.full(4097, 123456789, np.uint64), 33
    yield "u8_L28", rng.integers(0, 256, 9001).astype(np.uint8), 28
    for n in (1, 2, 63, 64, 65, 257):
        yield "tiny%d" % n, rng.integers(0, 1 << 26, n, dtype=np.uint64).astype(np.uint32), 26
    # full-width keys (the P_24 sort's bit range ends at the key's top bit), one
    # block and many
    yield "u32_L32_small", rng.integers(0, 1 << 32, 300, dtype=np.uint64).astype(np.uint32), 32
    yield "u64_L64_small", rng.integers(0, np.iinfo(np.uint64).max, 700, dtype=np.uint64), 64
    yield "u16_L32", rng.integers(0, 1 << 16, 40001).astype(np.uint16), 32
    yield "u32_L32_2e20", rng.integers(0, 1 << 32, (1 << 20) + 3, dtype=np.uint64).astype(np.uint32), 32


CASES = list(cases())


@pytest.mark.parametrize("method", [0, 1], ids=["top24_multilevel", "per_level"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_deep_levels(ctx, orc, case, method):
    """method 0: the top 24 levels by the multi-level passes, P_24 by a radix
    sort, the rest level by level; method 1: every level by the per-level steps."""
    name, s, L = CASES[case]
    print("deep case", name, s.size, L, method, flush=True)
    n = s.size
    ds = ctx.to_device(s)
    dw = ctx.empty(max(1, (n * L + 63) // 64), np.uint64)
    try:
        ctx.set_level_method(method)
        ctx.levels_build(ds, n, L, dw)
        got = dw.get()[:(n * L + 63) // 64]
        exp = closed_form(s, L)[:got.size]
        bad = np.nonzero(got != exp)[0]
        assert bad.size == 0, "%s: %d words differ, first at level %d" % (name, bad.size, bad[0] * 64 // n if bad.size else -1)
        if L <= 63:
            ow, _ = orc.levels_serial(s, L, with_nodes=False)
            assert np.array_equal(got, ow), name
    finally:
        ctx.set_level_method(0)
        dw.free()
        ds.free()


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("dt,L", [(np.uint32, 30), (np.uint64, 41)])
def test_deep_levels_bad_symbol(ctx, wsfm, method, dt, L):
    """A symbol >= 2^L is reported (by the multi-level count of the top levels,
    or by the first per-level step) and the context stays usable."""
    rng = np.random.default_rng(L)
    n = 50001
    s = rng.integers(0, 1 << L, n, dtype=np.uint64).astype(dt)
    s[[7, 30000]] = (1 << L) + 5
    print("deep bad symbol", np.dtype(dt).name, L, method, flush=True)
    ctx.set_level_method(method)
    try:
        with pytest.raises(wsfm.WtgError):
            ctx.levels(s, L)
        ctx.sync()
        good = s.copy()
        good[[7, 30000]] = 3
        w, _ = ctx.levels(good, L)
        assert np.array_equal(w[:(n * L + 63) // 64], closed_form(good, L)[:(n * L + 63) // 64])
    finally:
        ctx.set_level_method(0)


def test_deep_levels_reject_nodes(wsfm, ctx):
    s = np.arange(100, dtype=np.uint32)
    ds = ctx.to_device(s)
    dw = ctx.empty((100 * 30 + 63) // 64, np.uint64)
    dn = ctx.empty(5 * 64, np.uint64)
    try:
        with pytest.raises(wsfm.WtgError, match="24 levels"):
            ctx.levels_build(ds, 100, 30, dw, dn, 64)
    finally:
        dn.free()
        dw.free()
        ds.free()


@pytest.mark.parametrize("hi,dt", [(1 << 33, np.uint64), ((1 << 32) - 1, np.uint32), (1 << 50, np.uint64)])
def test_deep_wtint(ctx, hi, dt):
    """wt_int<> with max_level = hi(max) + 1 > 24: bits against the closed form,
    rank / access / select against brute force on the text."""
    rng = np.random.default_rng(hi & 0xFFFF)
    n = 30011
    base = rng.integers(0, hi, 500, dtype=np.uint64)   # repeated values, so select has work
    s = base[rng.integers(0, base.size, n)].astype(dt)
    print("deep wtint", hi, np.dtype(dt).name, flush=True)
    ds = ctx.to_device(s)
    wt = ctx.wtint_build(ds, n, 0)
    try:
        L = int(s.max()).bit_length()
        assert wt.L == L and L > 24 and wt.sigma == len(np.unique(s))
        assert np.array_equal(wt.words()[:(n * L + 63) // 64], closed_form(s, L)[:(n * L + 63) // 64])
        s64 = s.astype(np.uint64)
        ii = rng.integers(0, n, 2000).astype(np.uint64)
        assert np.array_equal(wt.access(ii), s64[ii.astype(np.int64)])
        cc = s64[rng.integers(0, n, 2000)]
        ri = rng.integers(0, n + 1, 2000).astype(np.uint64)
        exp = np.array([int((s64[:int(i)] == c).sum()) for i, c in zip(ri, cc)], np.uint64)
        assert np.array_equal(wt.rank(ri, cc), exp)
        occ = {int(v): np.nonzero(s64 == v)[0] for v in np.unique(cc)}
        js = np.array([1 + (k % max(1, occ[int(c)].size)) for k, c in enumerate(cc)], np.uint64)
        got = wt.select(js, cc)
        exp = np.array([occ[int(c)][int(j) - 1] for j, c in zip(js, cc)], np.uint64)
        assert np.array_equal(got, exp)
    finally:
        wt.free()
        ds.free()
