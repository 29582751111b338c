"""GPU parity of (a), the level-wise wavelet tree (wavelet/levelWT/WT.C:43-227),
against the oracle (itself pinned to the compiled reference) and the golden
fixtures written from the compiled reference.  Bit-exact words and nodes."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "levels_ref.npz")


def say(*a):
    print(*a, flush=True)


def golden_cases():
    z = np.load(GOLD, allow_pickle=False)
    names = sorted({k.split("__")[0] for k in z.files})
    return [(nm, z[nm + "__s"], int(z[nm + "__sigma"][0]), z[nm + "__words"], z[nm + "__reach"],
             z[nm + "__nodes"]) for nm in names]


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_wt_drop_in_matches_reference_golden(ctx, wsfm, case):
    """wtg_wt_host = the WT(s, n, sigma) drop-in (host buffers in and out)."""
    name, s, sigma, words, reach, nodes = case
    say("golden", name, s.size, sigma)
    w, nd = ctx.wt_host(s, sigma)
    assert np.array_equal(w, words)
    r = reach.astype(np.int64)
    assert np.array_equal(nd[r], nodes)
    # unreached slots are {0,0,NONE,NONE,NONE}
    mask = np.ones(nd.shape[0], bool)
    mask[r] = False
    if mask.any():
        assert (nd[mask, 0] == 0).all() and (nd[mask, 1] == 0).all()
        assert (nd[mask, 2:] == wsfm.NONE).all()


CASES = []
for dt in (np.uint8, np.uint16, np.uint32, np.uint64):
    for n in (1, 63, 64, 65, 1000, 4095, 4096, 4097, 16384, 100003):
        CASES.append((dt, n))


@pytest.mark.parametrize("dt,n", CASES, ids=lambda x: str(x))
def test_levels_vs_oracle_sizes(ctx, orc, dt, n):
    rng = np.random.default_rng(n)
    bits = np.dtype(dt).itemsize * 8
    for L in sorted({1, 2, 3, 7, min(bits, 8), min(bits, 13), min(bits, 20)}):
        s = rng.integers(0, 1 << L, n, dtype=np.uint64).astype(dt)
        say("levels", np.dtype(dt).name, n, L)
        w, nd = ctx.levels(s, L, with_nodes=L <= 16)
        ow, ond = orc.levels_serial(s, L, with_nodes=L <= 16)
        assert np.array_equal(w, ow), (n, L)
        if L <= 16:
            reached = np.nonzero(ond[:, 5])[0]
            assert np.array_equal(nd[reached], ond[reached, :5])


@pytest.mark.parametrize("seed", range(6))
def test_levels_skewed_and_sparse_alphabets(ctx, orc, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(20000, 300000))
    L = int(rng.integers(10, 25))
    kind = seed % 3
    if kind == 0:
        s = np.minimum(rng.geometric(1e-3, n) - 1, (1 << L) - 1)
    elif kind == 1:
        s = rng.choice(rng.integers(0, 1 << L, 37), n)
    else:
        s = np.where(rng.random(n) < 0.9, 5, rng.integers(0, 1 << L, n))
    s = s.astype(np.uint32)
    say("skewed", seed, n, L)
    w, _ = ctx.levels(s, L)
    ow, _ = orc.levels_serial(s, L, with_nodes=False)
    assert np.array_equal(w, ow)


def test_levels_empty_input(ctx, wsfm):
    w, nd = ctx.wt_host(np.zeros(0, np.uint32), 5)
    assert w.size == 0
    assert nd[0].tolist() == [0, 0, wsfm.NONE, wsfm.NONE, wsfm.NONE]


def test_remap_matches_driver(ctx, orc):
    rng = np.random.default_rng(3)
    s = rng.choice(np.array([5, 17, 1000, 65535, 3], np.uint32), 50000)
    d = ctx.to_device(s)
    sigma = ctx.remap(d)
    exp, esig = orc.remap(s)
    assert sigma == esig == 5
    assert np.array_equal(d.get(), exp)
    d.free()


def test_bad_symbol_reports_error(ctx, wsfm):
    s = np.array([1, 2, 300], np.uint32)  # 300 >= 2^8
    with pytest.raises(wsfm.WtgError):
        ctx.levels(s, 8)
    ctx.sync()


@pytest.mark.parametrize("dt,L", [(np.uint32, 8), (np.uint8, 5), (np.uint16, 11)])
def test_bad_symbol_in_full_tiles_reports_error(ctx, wsfm, dt, L):
    # invalid keys inside full tiles (the full-tile split kernel): the build
    # reports the error and every write stays inside the buffers; the context
    # stays usable
    rng = np.random.default_rng(L)
    n = 100003
    s = rng.integers(0, 1 << L, n).astype(dt)
    s[[0, 50000, 77777]] = (1 << L) + 3
    say("bad symbols in full tiles", np.dtype(dt).name, n, L)
    with pytest.raises(wsfm.WtgError):
        ctx.levels(s, L)
    ctx.sync()
    good = rng.integers(0, 1 << L, 5000).astype(dt)
    w, _ = ctx.levels(good, L, with_nodes=False)
    assert w.size == wsfm.lib().wtg_levels_words(good.size, L)


def test_levels_large_bit_exact(ctx, wsfm, orc):
    """n = 2^24 uint32, L = 20 (the wt_int shape at 1/16 of its full size):
    bit-exact against the oracle, plus per-level one counts."""
    n, L = 1 << 24, 20
    s = wsfm.gen_uniform(0x5EED2026, n, 1 << L)
    say("large", n, L)
    w, _ = ctx.levels(s, L)
    ow, _ = orc.levels_serial(s, L, with_nodes=False)
    assert np.array_equal(w, ow)
    bits = np.unpackbits(w.view(np.uint8), bitorder="little")[:n * L].reshape(L, n)
    for l in range(L):
        assert bits[l].sum() == int(((s >> (L - 1 - l)) & 1).sum())


def test_levels_geometric_twenty_levels(wsfm, orc, ctx):
    # the second wt_int input kind: geometric(1e-4) values capped at 2^20 - 1, 20 levels
    n, L = 1 << 21, 20
    s = wsfm.gen_geometric(21, n, 1e-4, (1 << L) - 1, np.uint32)
    print("geometric", n, L, flush=True)
    words, _ = ctx.levels(s, L)
    exp = orc.levels_closed(s, L)
    assert np.array_equal(words, exp)


class _At:
    """A device pointer into another array (an input that starts at any byte)."""

    def __init__(self, arr, first):
        self.dtype = np.dtype(arr.dtype)
        self.ptr = ctypes.c_void_p(arr.ptr.value + first * self.dtype.itemsize)


@pytest.mark.parametrize("dt,L", [(np.uint8, 8), (np.uint8, 5), (np.uint16, 13), (np.uint32, 20), (np.uint64, 22)])
def test_levels_unaligned_input(ctx, wsfm, orc, dt, L):
    """Inputs that start and end inside 16-byte words: the key loads cover the
    words around them, and symbols outside [first, first + n) (here invalid
    ones, >= 2^L) are neither counted nor reported."""
    rng = np.random.default_rng(L)
    N = 3 * 4096 + 77
    full = rng.integers(0, 1 << L, N).astype(dt)
    lw = wsfm.lib().wtg_levels_words
    for first, n in ((1, 5000), (3, 2048), (5, 9000), (13, 1), (7, 15), (1, N - 2)):
        s = full[first:first + n].copy()
        bad = full.copy()
        if (1 << L) + 1 <= np.iinfo(dt).max:  # invalid symbols right before and after the range
            bad[first - 1] = (1 << L) + 1
            if first + n < N:
                bad[first + n] = (1 << L) + 1
        say("unaligned", np.dtype(dt).name, L, first, n)
        db = ctx.to_device(bad)
        dw = ctx.empty(max(1, lw(n, L)), np.uint64)
        try:
            ctx.levels_build(_At(db, first), n, L, dw)
            got = dw.get()[:lw(n, L)]
            ow, _ = orc.levels_serial(s, L, with_nodes=False)
            assert np.array_equal(got, ow), (first, n)
        finally:
            dw.free()
            db.free()
