"""The allocation-free rank9 build (wtg_rank9_build_noalloc) against
wtg_rank9_build and the oracle's rank_support_v<1,1> directory
(rank/rank_support_v_par.hpp:69-123), on two calls in a row through the C ABI,
the second with a smaller vector (the shared scratch keeps the first call's
contents), and the size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(wsfm, ctx, dw, nbits, fill=0xA5A5A5A5A5A5A5A5):
    lib = wsfm.lib()
    D = ctx.to_device(np.full(lib.wtg_rank9_words(nbits), fill, np.uint64))
    wsfm._check(lib.wtg_rank9_build_noalloc(ctx.h, dw.ptr if dw is not None else None, int(nbits), D.ptr))
    out = D.get()
    D.free()
    return out


def test_noalloc_matches_rank9_build_twice(wsfm, ctx, orc):
    rng = np.random.default_rng(77)
    for nbits in ((1 << 27) + 4099, 100003, 64 * 512 * 8):
        print("rank9 noalloc", nbits, flush=True)
        nw = (nbits + 63) // 64
        w = rng.integers(0, 1 << 64, nw, dtype=np.uint64, endpoint=False)
        if nbits & 63:
            w[-1] &= np.uint64((1 << (nbits & 63)) - 1)
        # a sparse stretch and a full stretch so the 9-bit fields see 0 and 64
        w[nw // 3: nw // 3 + 700] = 0
        w[nw // 2: nw // 2 + 900] = np.uint64((1 << 64) - 1)
        dw = ctx.to_device(w)
        ref = ctx.rank9_build(dw, nbits)
        got = _noalloc(wsfm, ctx, dw, nbits)
        exp = ref.get()
        ref.free()
        dw.free()
        assert np.array_equal(got, exp)
        if nbits < (1 << 22):
            assert np.array_equal(got, orc.rank9_build(w, nbits))


def test_noalloc_empty_and_limit(wsfm, ctx):
    lib = wsfm.lib()
    got = _noalloc(wsfm, ctx, None, 0)
    assert np.array_equal(got, np.zeros(2, np.uint64))
    D = ctx.empty(2, np.uint64)
    rc = lib.wtg_rank9_build_noalloc(ctx.h, D.ptr, (1 << 35) + 1, D.ptr)
    D.free()
    assert rc == -4  # WTG_ERR_UNSUPPORTED, rejected before any launch
    assert b"WTG_RANK9_NOALLOC_MAX_BITS" in lib.wtg_last_error()
