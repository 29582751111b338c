"""Host side of the FM-index workload (no GPU): the SA-IS suffix array against
a comparison sort (the order divsufsort produces, fm-index/par/WT.C:35-36), the
BWT of fm-index/par/WT.C:38-41, and the DNA text / pattern generators."""
import numpy as np
import pytest


@pytest.mark.parametrize("k", [1, 2, 3, 4, 26])
def test_suffix_array_matches_sort(wsfm, orc, k):
    rng = np.random.default_rng(k)
    for trial in range(120):
        n = int(rng.integers(1, 400))
        t = (rng.integers(0, k, n) + 97).astype(np.uint8)
        if trial % 3 == 0:
            t[-1] = 0                     # with a unique smallest last byte
        if trial % 7 == 1:
            t[:] = 97                     # a run: every suffix a prefix of the next longer
        sa = wsfm.suffix_array(t)
        assert np.array_equal(sa, orc.sa_brute(t)), (k, trial)


def test_suffix_array_periodic_and_dna(wsfm, orc):
    t = np.frombuffer(b"abaababaabaababaababa" * 40, np.uint8).copy()
    assert np.array_equal(wsfm.suffix_array(t), orc.sa_brute(t))
    t = wsfm.gen_dna(1, 30000)
    sa = wsfm.suffix_array(t)
    assert np.array_equal(sa, orc.sa_brute(t))
    bwt = wsfm.bwt_from_sa(t, sa)
    assert np.array_equal(bwt, orc.bwt_from_sa(t, sa))
    # the BWT is a permutation of the text and, with the sentinel, inverts to it
    assert np.array_equal(np.sort(bwt), np.sort(t))
    assert bwt[np.nonzero(sa == 0)[0][0]] == t[-1]


def test_empty_and_bad_args(wsfm):
    assert wsfm.suffix_array(np.zeros(0, np.uint8)).size == 0
    with pytest.raises(wsfm.WtgError):
        wsfm.gen_dna(1, 1)


def test_gen_dna_properties(wsfm):
    n = 1 << 20
    a = wsfm.gen_dna(5, n)
    assert np.array_equal(a, wsfm.gen_dna(5, n))
    assert not np.array_equal(a, wsfm.gen_dna(6, n))
    assert a[-1] == 0 and (a[:-1] != 0).all()
    cnt = np.bincount(a, minlength=256)
    assert cnt[[65, 67, 71, 78, 84]].sum() == n - 1
    assert 0.0005 < cnt[78] / n < 0.0012          # ~0.1% N (draws may repeat a position)


def test_gen_patterns(wsfm, orc):
    t = wsfm.gen_dna(7, 50000)
    p = wsfm.gen_patterns(8, t, 20, 400)
    assert p.shape == (400, 20)
    assert np.array_equal(p, wsfm.gen_patterns(8, t, 20, 400))
    sa = wsfm.suffix_array(t)
    cnt = orc.sa_count(t, sa, p)
    assert (cnt[0::2] >= 1).all()                  # exact substrings occur
    assert (p != 0).all()                          # never the sentinel


def test_oracle_fm_count_pinned(wsfm, orc):
    # backward search (oracle) = SA binary search = brute-force substring count
    t = wsfm.gen_dna(9, 4000)
    sa = wsfm.suffix_array(t)
    bwt = wsfm.bwt_from_sa(t, sa)
    p = np.concatenate([wsfm.gen_patterns(10, t, 4, 300), wsfm.gen_patterns(11, t, 1, 50)[:, :1].repeat(4, 1)])
    fm = orc.fm_count(bwt, p)
    assert np.array_equal(fm, orc.sa_count(t, sa, p))
    s = bytes(t)
    for q in range(0, p.shape[0], 7):
        pat = bytes(p[q])
        brute = sum(1 for i in range(len(s) - len(pat) + 1) if s[i:i + len(pat)] == pat)
        assert fm[q] == brute
