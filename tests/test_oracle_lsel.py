"""CPU checks of the learned-select oracle and workload (no GPU): the closed-form
spline gradient against central finite differences of the loss, the loss at
the least-squares optimum, and the C5 bitvector generator's densities."""
import numpy as np
import pytest


def random_samples(rng, ones, stride, d):
    # positions of the (t+1)*stride-th ones of a Bernoulli(d) bitvector
    gaps = rng.geometric(d, ones)
    pos = np.cumsum(gaps) - 1
    return pos[stride - 1::stride][: ones // stride].astype(np.float64)


@pytest.mark.parametrize("spacing,stride", [(64, 16), (256, 16), (128, 32)])
def test_spline_gradient_finite_difference(orc, spacing, stride):
    rng = np.random.default_rng(spacing + stride)
    ones = 5000
    pos = random_samples(rng, ones, stride, 0.05)
    ns = pos.size
    nk = (ns * stride) // spacing + 2
    alpha = (pos[-1] - pos[0]) / ((ns - 1) * stride)
    beta = pos[0] - alpha * stride
    y = rng.normal(0, 50, nk)
    loss, g = orc.spline_loss_grad(y, spacing, stride, alpha, beta, pos)
    h = 1e-3
    for k in range(0, nk, max(1, nk // 25)):
        yp, ym = y.copy(), y.copy()
        yp[k] += h
        ym[k] -= h
        fd = (orc.spline_loss_grad(yp, spacing, stride, alpha, beta, pos)[0] -
              orc.spline_loss_grad(ym, spacing, stride, alpha, beta, pos)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-7
    # prediction restates the loss's residual
    j = (np.arange(ns) + 1) * stride
    e = orc.spline_predict(y, spacing, alpha, beta, j) - pos
    assert abs(loss - float((e * e).mean())) < 1e-9 * max(1, loss)


def test_spline_least_squares_optimum(orc):
    # the gradient vanishes at the least-squares knots (solved directly)
    rng = np.random.default_rng(3)
    stride, spacing = 16, 128
    pos = random_samples(rng, 4000, stride, 0.1)
    ns = pos.size
    nk = (ns * stride) // spacing + 2
    j = (np.arange(ns) + 1) * stride
    A = np.zeros((ns, nk))
    k = j // spacing
    f = (j - k * spacing) / spacing
    A[np.arange(ns), k] = 1 - f
    A[np.arange(ns), k + 1] = f
    yls, *_ = np.linalg.lstsq(A, pos, rcond=None)
    loss, g = orc.spline_loss_grad(yls, spacing, stride, 0.0, 0.0, pos)
    assert np.abs(g).max() < 1e-6 * max(1, np.abs(pos).max())
    assert loss <= orc.spline_loss_grad(yls + rng.normal(0, 1, nk), spacing, stride, 0.0, 0.0, pos)[0]


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
