"""Learned select on the GPU: the training kernels' loss and gradient against
the float64 oracle (tolerance: relative 2e-3 of the largest gradient entry,
float32 accumulation over at most spacing/stride samples per knot interval),
exact answers (learned == plain sampled select == brute force), the
exhaustively measured error window, and what training gains over the
untrained knots."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits_of(w, nb):
    return np.unpackbits(w.view(np.uint8), bitorder="little")[:nb]


@pytest.mark.parametrize("spacing,stride,d", [(64, 16, 0.05), (512, 16, 0.3), (256, 32, 0.01)])
def test_loss_grad_matches_oracle(wsfm, orc, ctx, spacing, stride, d):
    rng = np.random.default_rng(spacing)
    ones = 200000
    pos = (np.cumsum(rng.geometric(d, ones)) - 1)[stride - 1::stride][: ones // stride].astype(np.uint64)
    ns = pos.size
    nk = (ns * stride) // spacing + 2
    alpha = (float(pos[-1]) - float(pos[0])) / ((ns - 1) * stride)
    beta = float(pos[0]) - alpha * stride
    y = rng.normal(0, 30, nk).astype(np.float32)
    loss, g = wsfm.lsel_loss_grad(ctx, y, spacing, stride, alpha, beta, pos)
    oloss, og = orc.spline_loss_grad(y.astype(np.float64), spacing, stride, alpha, beta, pos)
    assert abs(loss - oloss) <= 2e-3 * oloss
    assert np.abs(g - og).max() <= 2e-3 * np.abs(og).max()


def make_vectors(wsfm, ctx, nvec, nb, seed):
    w, dens = wsfm.gen_select_vectors(seed, nvec, nb)
    nw = w.shape[1]
    dv = []
    for b in range(nvec):
        padded = np.zeros(nw + 1, np.uint64)
        padded[:nw] = w[b]
        dv.append(ctx.to_device(padded))
    return w, dens, dv


def test_learned_select_exact(wsfm, orc, ctx):
    nvec, nb = 8, 1 << 21
    print("lsel exact", nvec, nb, flush=True)
    w, dens, dv = make_vectors(wsfm, ctx, nvec, nb, 21)
    ls = wsfm.LearnedSelect(ctx, dv, [nb] * nvec)
    try:
        rng = np.random.default_rng(0)
        learned_any = False
        for b in range(nvec):
            info = ls.info(b)
            ones = np.flatnonzero(bits_of(w[b], nb))
            assert info["ones"] == ones.size
            if info["spacing"]:
                learned_any = True
                assert info["eps"] <= 512
                assert info["nknots"] == ones.size // info["spacing"] + 2
            # every rank for two vectors, a sample for the others, plus the edges
            j = np.arange(1, ones.size + 1) if b < 2 else rng.integers(1, ones.size + 1, 20000)
            j = np.concatenate([j, [0, ones.size, ones.size + 1]]).astype(np.uint64)
            vec = np.full(j.size, b, np.uint32)
            got = ls.select(vec, j)
            plain = ls.select(vec, j, learned=False)
            assert np.array_equal(got, plain)
            exp = np.full(j.size, orc.NONE, np.uint64)
            ok = (j >= 1) & (j <= ones.size)
            exp[ok] = ones[j[ok].astype(np.int64) - 1]
            assert np.array_equal(got, exp)
        assert learned_any
        # mixed batch across vectors, out-of-range vector ids
        vv = rng.integers(0, nvec + 1, 50000).astype(np.uint32)
        jj = rng.integers(1, 3000, 50000).astype(np.uint64)
        assert np.array_equal(ls.select(vv, jj), ls.select(vv, jj, learned=False))
    finally:
        ls.free()
        for d in dv:
            d.free()


def test_training_gain(wsfm, ctx):
    # dense Bernoulli and run vectors at one fixed spacing: Adam lowers the
    # training loss below the untrained (knot-sample) initialisation
    nvec, nb = 4, 1 << 22
    w, dens, dv = make_vectors(wsfm, ctx, nvec, nb, 33)
    p = wsfm.LselParams(16, 1024, 1024, 300, 4.0, 0xFFFFFFFF)
    ls = wsfm.LearnedSelect(ctx, dv, [nb] * nvec, p)
    try:
        for b in range(nvec):
            info = ls.info(b)
            print("vec", b, "density %.4f" % dens[b], info, flush=True)
            assert info["spacing"] == 1024
            assert info["loss_final"] < info["loss_init"]
            assert info["eps"] <= info["eps_init"] * 1.5 + 64
    finally:
        ls.free()
        for d in dv:
            d.free()


def test_bad_params(wsfm, ctx):
    d = ctx.to_device(np.zeros(3, np.uint64))
    with pytest.raises(wsfm.WtgError):
        wsfm.LearnedSelect(ctx, [d], [128], wsfm.LselParams(16, 64, 16, 10, 1.0, 512))
    d.free()
