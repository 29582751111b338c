"""Learned select on the GPU: the training kernels' loss and gradient against
the float64 oracle (tolerance: relative 2e-3 of the largest gradient entry,
float32 accumulation over the samples of one knot interval), the knot placement
against the oracle's restatement (knot arguments identical, values within
float32 rounding), the exhaustively measured error window against the same
window recomputed on the host from the downloaded model, exact answers
(learned == plain sampled select == brute force, edge sizes included), and what
training does to the placed knots."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits_of(w, nb):
    return np.unpackbits(w.view(np.uint8), bitorder="little")[:nb]


@pytest.mark.parametrize("spacing,stride,d", [(64, 16, 0.05), (512, 16, 0.3), (256, 32, 0.01), (0, 16, 0.02)])
def test_loss_grad_matches_oracle(wsfm, orc, ctx, spacing, stride, d):
    rng = np.random.default_rng(spacing + 1)
    ones = 200000
    pos = (np.cumsum(rng.geometric(d, ones)) - 1)[stride - 1::stride][: ones // stride].astype(np.uint64)
    ns = pos.size
    if spacing:
        J = np.arange((ns * stride) // spacing + 2, dtype=np.uint64) * spacing
    else:  # irregular knots
        J = np.concatenate([[1], np.sort(rng.choice(np.arange(2, ns * stride), 3000, replace=False)),
                            [ns * stride]]).astype(np.uint64)
    nk = J.size
    alpha = (float(pos[-1]) - float(pos[0])) / ((ns - 1) * stride)
    beta = float(pos[0]) - alpha * stride
    y = rng.normal(0, 30, nk).astype(np.float32)
    loss, g = wsfm.lsel_loss_grad(ctx, J, y, stride, alpha, beta, pos)
    oloss, og = orc.spline_loss_grad(J.astype(np.float64), y.astype(np.float64), stride, alpha, beta, pos)
    assert abs(loss - oloss) <= 2e-3 * oloss
    assert np.abs(g - og).max() <= 2e-3 * np.abs(og).max()


def make_vectors(ctx, arrays):
    dv = []
    for w in arrays:
        padded = np.zeros(w.size + 1, np.uint64)
        padded[:w.size] = w
        dv.append(ctx.to_device(padded))
    return dv


def pack(bits):
    nb = bits.size
    w = np.zeros((nb + 63) // 64 * 8, np.uint8)
    w[: (nb + 7) // 8] = np.packbits(bits, bitorder="little")
    return w.view(np.uint64)


def host_eps(orc, knots, alpha, beta, ones_pos):
    """max |round(pred(j)) - select(j)| from the downloaded model (f32
    interpolation as on the device; the device may fuse one multiply-add)."""
    J = knots["j"].astype(np.float64)
    y = knots["y"]
    j = np.arange(1, ones_pos.size + 1, dtype=np.float64)
    k = orc.spline_interval(J, j)
    f = ((j - J[k]).astype(np.float32) / (J[k + 1] - J[k]).astype(np.float32)).astype(np.float32)
    s = (np.float32(1) - f) * y[k] + f * y[k + 1]
    p = np.rint(orc.model_line(alpha, beta, j) + s.astype(np.float64))
    return int(np.abs(p - ones_pos).max())


def test_placement_matches_oracle(wsfm, orc, ctx):
    rng = np.random.default_rng(11)
    nb = (1 << 18) + 77
    vecs = [(rng.random(nb) < 0.02), (rng.random(nb) < 0.3)]
    for d in (0.05, 0.4):   # run-structured: 1-runs mean 64
        bits = np.zeros(nb, bool)
        p, one = 0, False
        while p < nb:
            ln = int(rng.geometric(1 / 64 if one else d / (64 * (1 - d))))
            bits[p:p + ln] = one
            p += ln
            one = not one
        vecs.append(bits)
    # very sparse (64-word windows with no one or a few), and one long run of
    # ones between sparse stretches (windows of 64 ones per word)
    vecs.append(rng.random(nb) < 0.0015)
    dense = rng.random(nb) < 0.01
    dense[70000:170000] = True
    vecs.append(dense)
    words = [pack(b.astype(np.uint8)) for b in vecs]
    dv = make_vectors(ctx, words)
    tol, gap = 64, 1024
    ls = wsfm.LearnedSelect(ctx, dv, [nb] * len(vecs), wsfm.LselParams(16, gap, tol, 0, 4.0, 512))
    try:
        for b, bits in enumerate(vecs):
            print("placement vector", b, flush=True)
            ones = np.flatnonzero(bits).astype(np.float64)
            info = ls.info(b)
            assert info["learned"] == 1 and info["trained"] == 1
            knots, lut, lut_log, alpha, beta = ls.model(b)
            J, V = orc.place_knots(ones, float(tol), gap)
            assert np.array_equal(knots["j"].astype(np.uint64), J)
            res = (V - orc.model_line(alpha, beta, J.astype(np.float64))).astype(np.float32)
            assert np.allclose(knots["y"], res, rtol=0, atol=1e-3 + 1e-6 * np.abs(res).max())
            # lookup table: entry i = the interval holding argument i << lut_log
            i = np.arange(lut.size, dtype=np.float64) * (1 << lut_log)
            assert np.array_equal(lut, orc.spline_interval(J.astype(np.float64), np.maximum(i, 1)))
            e = host_eps(orc, knots, alpha, beta, ones)
            assert abs(info["eps"] - e) <= 1 and info["eps"] <= tol + 1
            assert info["eps"] == info["eps_init"]
    finally:
        ls.free()
        for d in dv:
            d.free()


def test_learned_select_exact(wsfm, orc, ctx):
    nvec, nb = 8, 1 << 21
    print("lsel exact", nvec, nb, flush=True)
    w, dens = wsfm.gen_select_vectors(21, nvec, nb)
    arrays = [w[b] for b in range(nvec)]
    sizes = [nb] * nvec
    # edge sizes: partial last rank9 block, all ones, a single one, no ones, few ones
    rng = np.random.default_rng(5)
    extra = [(rng.random(nb + 37) < 0.1), np.ones(5000, bool), np.zeros(1000, bool),
             np.zeros(4099, bool), (rng.random(700) < 0.05)]
    extra[2][999] = True
    for bits in extra:
        arrays.append(pack(bits.astype(np.uint8)))
        sizes.append(bits.size)
    dv = make_vectors(ctx, arrays)
    ls = wsfm.LearnedSelect(ctx, dv, sizes)
    try:
        for b in range(len(arrays)):
            info = ls.info(b)
            ones = np.flatnonzero(bits_of(arrays[b], sizes[b]))
            assert info["ones"] == ones.size
            if ones.size >= 32:   # two training samples at stride 16
                assert info["learned"] == 1 and info["eps"] <= 512, (b, info)
            else:
                assert info["learned"] == 0
            # every rank for three vectors, a sample for the others, plus the edges
            j = np.arange(1, ones.size + 1) if b < 3 or b >= nvec else rng.integers(1, ones.size + 1, 20000)
            j = np.concatenate([j, [0, ones.size, ones.size + 1]]).astype(np.uint64)
            vec = np.full(j.size, b, np.uint32)
            got = ls.select(vec, j)
            plain = ls.select(vec, j, learned=False)
            assert np.array_equal(got, plain), b
            exp = np.full(j.size, orc.NONE, np.uint64)
            ok = (j >= 1) & (j <= ones.size)
            exp[ok] = ones[j[ok].astype(np.int64) - 1]
            assert np.array_equal(got, exp), b
        # mixed batch across vectors, out-of-range vector ids
        vv = rng.integers(0, len(arrays) + 1, 50000).astype(np.uint32)
        jj = rng.integers(1, 3000, 50000).astype(np.uint64)
        assert np.array_equal(ls.select(vv, jj), ls.select(vv, jj, learned=False))
    finally:
        ls.free()
        for d in dv:
            d.free()


def test_training_effect(wsfm, ctx):
    # Bernoulli and run vectors: Adam lowers the training loss of the placed
    # knots on every vector; the kept model stays within eps_max
    nvec, nb = 4, 1 << 22
    w, dens = wsfm.gen_select_vectors(33, nvec, nb)
    dv = make_vectors(ctx, [w[b] for b in range(nvec)])
    ls = wsfm.LearnedSelect(ctx, dv, [nb] * nvec, wsfm.LselParams(16, 4096, 256, 300, 4.0, 512))
    try:
        for b in range(nvec):
            info = ls.info(b)
            print("vec", b, "density %.4f" % dens[b], info, flush=True)
            assert info["learned"] == 1
            assert info["loss_final"] < info["loss_init"]
            assert info["eps_init"] <= 257 and info["eps"] <= 512
    finally:
        ls.free()
        for d in dv:
            d.free()


def test_bad_params(wsfm, ctx):
    d = ctx.to_device(np.zeros(3, np.uint64))
    with pytest.raises(wsfm.WtgError):   # tol must stay below eps_max - 1
        wsfm.LearnedSelect(ctx, [d], [128], wsfm.LselParams(16, 64, 600, 10, 1.0, 512))
    with pytest.raises(wsfm.WtgError):   # eps_max above the three-block window
        wsfm.LearnedSelect(ctx, [d], [128], wsfm.LselParams(16, 64, 100, 10, 1.0, 1024))
    with pytest.raises(wsfm.WtgError):   # max_gap not a power of two
        wsfm.LearnedSelect(ctx, [d], [128], wsfm.LselParams(16, 100, 100, 10, 1.0, 512))
    d.free()


@pytest.mark.parametrize("tol,gap", [(480, 4096), (200, 16384)])
def test_placement_matches_oracle_wide_chunks(wsfm, orc, ctx, tol, gap):
    """The default chunk (4096 arguments) and a wider one, tolerances as used."""
    rng = np.random.default_rng(tol)
    nb = (1 << 19) + 5
    vecs = [rng.random(nb) < 0.05, rng.random(nb) < 0.6]
    words = [pack(b.astype(np.uint8)) for b in vecs]
    dv = make_vectors(ctx, words)
    ls = wsfm.LearnedSelect(ctx, dv, [nb] * len(vecs), wsfm.LselParams(16, gap, tol, 0, 4.0, 512))
    try:
        for b, bits in enumerate(vecs):
            print("placement vector", b, tol, gap, flush=True)
            ones = np.flatnonzero(bits).astype(np.float64)
            knots, lut, lut_log, alpha, beta = ls.model(b)
            J, V = orc.place_knots(ones, float(tol), gap)
            assert np.array_equal(knots["j"].astype(np.uint64), J)
            res = (V - orc.model_line(alpha, beta, J.astype(np.float64))).astype(np.float32)
            assert np.allclose(knots["y"], res, rtol=0, atol=1e-3 + 1e-6 * np.abs(res).max())
            assert ls.info(b)["eps"] <= tol + 1
    finally:
        ls.free()
        for d in dv:
            d.free()


def test_training_gain(wsfm, ctx):
    """What training gains over the placed model, on the same vectors and
    queries: Adam lowers the mean squared error of every trained model, while
    the query cost stays the same: once eps <= 512 the correction reads the
    same three-block window, so the 64-byte sectors per query agree within
    alignment noise.  Answers are identical on the plain, placed and trained
    paths.  (The gain over the plain sampled select is its query rate, which
    the C5 bench measures at full size.)"""
    nvec, nb = 8, 1 << 22
    w, dens = wsfm.gen_select_vectors(41, nvec, nb)
    dv = make_vectors(ctx, [w[b] for b in range(nvec)])
    placed = wsfm.LearnedSelect(ctx, dv, [nb] * nvec, wsfm.LselParams(16, 4096, 480, 0, 4.0, 512))
    trained = wsfm.LearnedSelect(ctx, dv, [nb] * nvec, wsfm.LselParams(16, 4096, 480, 200, 4.0, 512))
    rng = np.random.default_rng(3)
    try:
        vec, j = [], []
        for b in range(nvec):
            ones = trained.info(b)["ones"]
            vec.append(np.full(50000, b, np.uint32))
            j.append(rng.integers(1, ones + 1, 50000).astype(np.uint64))
        dvec, dj = ctx.to_device(np.concatenate(vec)), ctx.to_device(np.concatenate(j))
        outs, secs = {}, {}
        for name, ls, learned in (("plain", placed, False), ("placed", placed, True), ("trained", trained, True)):
            do = ctx.empty(dj.size, np.uint64)
            ls.query(dvec, dj, do, learned)
            outs[name] = do.get()
            ds = ls.sectors(dvec, dj, do, learned)
            secs[name] = float(ds.get().mean())
            ds.free()
            do.free()
        ntrained = sum(trained.info(b)["trained"] for b in range(nvec))
        print("sectors per query", secs, "trained vectors", ntrained, flush=True)
        assert np.array_equal(outs["placed"], outs["plain"])
        assert np.array_equal(outs["trained"], outs["plain"])
        assert abs(secs["trained"] - secs["placed"]) < 0.05
        for b in range(nvec):
            info = trained.info(b)
            if info["trained"]:
                assert info["loss_final"] < info["loss_init"] and info["eps"] <= 512
        dvec.free()
        dj.free()
    finally:
        placed.free()
        trained.free()
        for d in dv:
            d.free()
