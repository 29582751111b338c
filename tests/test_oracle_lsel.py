"""CPU checks of the learned-select oracle and workload (no GPU): the closed-form
spline gradient against central finite differences of the loss (uniform and
irregular knots), the vanishing gradient at the least-squares optimum, the
greedy knot placement against its defining properties (every one within tol,
every segment maximal, chunk edges) and closed-form cases, and the C5
bitvector generator's densities."""
import numpy as np
import pytest


def random_samples(rng, ones, stride, d):
    # positions of the (t+1)*stride-th ones of a Bernoulli(d) bitvector
    gaps = rng.geometric(d, ones)
    pos = np.cumsum(gaps) - 1
    return pos[stride - 1::stride][: ones // stride].astype(np.float64)


def irregular_knots(rng, last, nk):
    inner = np.sort(rng.choice(np.arange(2, last), nk - 2, replace=False))
    return np.concatenate([[1], inner, [last]]).astype(np.float64)


@pytest.mark.parametrize("spacing,stride,irregular", [(64, 16, False), (256, 16, False), (128, 32, False),
                                                      (0, 16, True), (0, 1, True)])
def test_spline_gradient_finite_difference(orc, spacing, stride, irregular):
    rng = np.random.default_rng(spacing + stride + 7 * irregular)
    ones = 5000
    pos = random_samples(rng, ones, stride, 0.05)
    ns = pos.size
    if irregular:
        J = irregular_knots(rng, ns * stride, 60)
    else:
        J = np.arange((ns * stride) // spacing + 2, dtype=np.float64) * spacing
    nk = J.size
    alpha = (pos[-1] - pos[0]) / ((ns - 1) * stride)
    beta = pos[0] - alpha * stride
    y = rng.normal(0, 50, nk)
    loss, g = orc.spline_loss_grad(J, y, stride, alpha, beta, pos)
    h = 1e-3
    for k in range(0, nk, max(1, nk // 25)):
        yp, ym = y.copy(), y.copy()
        yp[k] += h
        ym[k] -= h
        fd = (orc.spline_loss_grad(J, yp, stride, alpha, beta, pos)[0] -
              orc.spline_loss_grad(J, ym, stride, alpha, beta, pos)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-7
    # prediction restates the loss's residual
    j = (np.arange(ns) + 1) * stride
    e = orc.spline_predict(J, y, alpha, beta, j) - pos
    assert abs(loss - float((e * e).mean())) < 1e-9 * max(1, loss)


def test_spline_least_squares_optimum(orc):
    # the gradient vanishes at the least-squares knots (solved directly)
    rng = np.random.default_rng(3)
    stride = 16
    pos = random_samples(rng, 4000, stride, 0.1)
    ns = pos.size
    J = irregular_knots(rng, ns * stride, 40)
    nk = J.size
    j = (np.arange(ns) + 1) * stride
    A = np.zeros((ns, nk))
    k = orc.spline_interval(J, j)
    f = (j - J[k]) / (J[k + 1] - J[k])
    A[np.arange(ns), k] = 1 - f
    A[np.arange(ns), k + 1] = f
    yls, *_ = np.linalg.lstsq(A, pos, rcond=None)
    loss, g = orc.spline_loss_grad(J, yls, stride, 0.0, 0.0, pos)
    assert np.abs(g).max() < 1e-6 * max(1, np.abs(pos).max())
    assert loss <= orc.spline_loss_grad(J, yls + rng.normal(0, 1, nk), stride, 0.0, 0.0, pos)[0]


def interp(J, V, j):
    k = np.clip(np.searchsorted(J, j, side="right") - 1, 0, J.size - 2)
    f = (j - J[k]) / (J[k + 1] - J[k])
    return (1 - f) * V[k] + f * V[k + 1]


def cone_feasible(x0, y0, xs, ps, tol):
    lo = ((ps - tol - y0) / (xs - x0)).max()
    hi = ((ps + tol - y0) / (xs - x0)).min()
    return lo <= hi


def run_vector(rng, nbits, d):
    bits = np.zeros(nbits, np.uint8)
    p, one = 0, False
    while p < nbits:
        ln = int(rng.geometric(1 / 64 if one else d / (64 * (1 - d)) if d < 1 else 1))
        if one:
            bits[p:p + ln] = 1
        p += ln
        one = not one
    return bits


@pytest.mark.parametrize("kind,d,tol,gap", [("bern", 0.02, 64, 512), ("bern", 0.3, 16, 256),
                                            ("runs", 0.05, 64, 1024), ("runs", 0.4, 256, 4096)])
def test_place_knots_properties(orc, kind, d, tol, gap):
    rng = np.random.default_rng(int(d * 1000) + tol)
    nbits = 1 << 17
    bits = (rng.random(nbits) < d).astype(np.uint8) if kind == "bern" else run_vector(rng, nbits, d)
    ones = np.flatnonzero(bits).astype(np.float64)
    m = ones.size
    J, V = orc.place_knots(ones, tol, gap)
    Jf = J.astype(np.float64)
    assert J[0] == 1 and J[-1] == m and (np.diff(J.astype(np.int64)) > 0).all()
    # chunk edges are knots
    starts = np.arange(0, m, gap) + 1
    ends = np.minimum(np.arange(1, (m + gap - 1) // gap + 1) * gap, m)
    assert np.isin(starts, J).all() and np.isin(ends, J).all()
    # chunk-start knots sit on their one exactly
    assert np.array_equal(V[np.isin(J, starts)], ones[starts - 1])
    # every one within tol of the interpolation
    j = np.arange(1, m + 1, dtype=np.float64)
    err = np.abs(interp(Jf, V, j) - ones)
    assert err.max() <= tol * (1 + 1e-12) + 1e-6
    # every segment that a chunk edge does not end is maximal: one more argument
    # empties the cone of lines from its first knot
    edge = set(ends.tolist())
    checked = 0
    for k in range(J.size - 1):
        a, b = int(J[k]), int(J[k + 1])
        if b in edge or a in edge:
            continue
        xs = np.arange(a + 1, b + 2, dtype=np.float64)
        assert not cone_feasible(float(a), V[k], xs, ones[a:b + 1], tol), (k, a, b)
        checked += 1
    assert checked > 0 or J.size <= 2 * len(starts)
    # a gap wider than 4*tol + 1 between two runs of at least two ones: no line
    # within tol spans it and a neighbour, so both its ends are knots
    dif = np.diff(ones)
    i = np.flatnonzero(dif > 4 * tol + 1) + 1      # ones[i-1] -> ones[i] is the gap
    i = i[(i >= 2) & (i + 1 < m)]
    i = i[(dif[i - 2] == 1) & (dif[i] == 1)]
    assert np.isin(i, J).all() and np.isin(i + 1, J).all()   # arguments i and i+1
    if kind == "runs" and 64 * (1 - d) / d > 8 * tol:   # mean gap far above the window
        assert i.size > 10


@pytest.mark.parametrize("step,gap", [(1, 64), (10, 100), (3, 4096)])
def test_place_knots_straight_line(orc, step, gap):
    # ones every `step` bits: pos(j) = step*(j-1) is a line, so each chunk is one
    # segment: knots exactly at the chunk starts and ends, on the ones
    m = 1000
    ones = step * np.arange(m, dtype=np.float64)
    J, V = orc.place_knots(ones, 8, gap)
    starts = np.arange(0, m, gap) + 1
    ends = np.minimum(np.arange(1, (m + gap - 1) // gap + 1) * gap, m)
    exp = np.unique(np.concatenate([starts, ends]))
    assert np.array_equal(J, exp)
    assert np.allclose(V, ones[exp - 1], atol=1e-9)


def test_select_vector_generator(wsfm):
    nb = 1 << 20
    w, d = wsfm.gen_select_vectors(5, 8, nb)
    w2, d2 = wsfm.gen_select_vectors(5, 8, nb)
    assert np.array_equal(w, w2) and np.array_equal(d, d2)
    assert ((d >= 0.01) & (d <= 0.5)).all()
    for b in range(8):
        bits = np.unpackbits(w[b].view(np.uint8), bitorder="little")[:nb]
        dens = bits.mean()
        assert abs(dens - d[b]) < 0.35 * d[b] + 0.002, (b, dens, d[b])
        if b % 2 == 1:   # run structure: mean 1-run length ~64
            edges = np.flatnonzero(np.diff(bits.astype(np.int8)) == -1).size
            assert edges > 0 and 30 < bits.sum() / edges < 130
