"""Pins the oracle's rank9 directory and select_support_mcl content.

The reference rank/select code needs sdsl-lite (absent), so it cannot be run.
The oracle (C) is checked against:
  * an independent vectorised numpy restatement of the directory fields
    (rank/rank_support_v_par.hpp:78-113, rankSerial/rank_support_v.hpp:76-106),
  * brute-force prefix popcounts for every index,
  * an independent pure-Python restatement of init_slow
    (select/select_support_mcl_par.hpp:227-276),
  * brute-force argument positions for every select(i).
"""
import numpy as np
import pytest


def bits_of(words, nbits):
    b = np.unpackbits(words.view(np.uint8), bitorder="little")
    return b[:nbits].astype(np.int64)


def rand_words(rng, nbits, density=0.5, runs=False):
    nw = (nbits + 63) // 64
    if runs:
        bits = np.zeros(nw * 64, np.uint8)
        pos = 0
        while pos < nbits:
            one = int(rng.geometric(1 / 64.0))
            zero = int(rng.geometric(density / (64 * (1 - density) + 1e-9))) if density < 1 else 0
            bits[pos:pos + one] = 1
            pos += one + zero
    else:
        bits = (rng.random(nw * 64) < density).astype(np.uint8)
    bits[nbits:] = 0
    return np.packbits(bits, bitorder="little").view(np.uint64).copy()


def numpy_rank9(words, nbits):
    """Independent restatement: per-word popcounts -> entries of 8 words."""
    if nbits == 0:
        return np.zeros(2, np.uint64)
    cap = (nbits + 63) // 64 * 64
    W, S = cap // 64, cap // 512
    pc = np.array([bin(int(x)).count("1") for x in words[:W]], np.int64)
    pc = np.concatenate([pc, np.zeros(8 * (S + 1) - W, np.int64)])
    D = np.zeros(2 * (S + 1), np.uint64)
    run = 0
    for b in range(S + 1):
        D[2 * b] = run
        m = max(0, min(8, W - 8 * b))
        c = np.concatenate([[0], np.cumsum(pc[8 * b:8 * b + 8])])
        f = 0
        for j in range(1, 8):
            if j < m or (j == m and m < 8):
                f |= int(c[j]) << (63 - 9 * j)
        if m == 0:
            f = 0
        D[2 * b + 1] = f
        run += int(c[8])
    return D


SIZES = [0, 1, 63, 64, 65, 511, 512, 513, 1000, 4096, 4097 * 8, 100000]


@pytest.mark.parametrize("nbits", SIZES)
def test_rank9_directory_vs_numpy(orc, nbits):
    rng = np.random.default_rng(nbits + 1)
    w = rand_words(rng, max(nbits, 1), 0.4)[:(nbits + 63) // 64]
    assert np.array_equal(orc.rank9_build(w, nbits), numpy_rank9(w, nbits))


@pytest.mark.parametrize("nbits", [1, 64, 513, 5000, 70000])
def test_rank9_every_index(orc, nbits):
    rng = np.random.default_rng(nbits)
    w = rand_words(rng, nbits, 0.3)
    D = orc.rank9_build(w, nbits)
    pref = np.concatenate([[0], np.cumsum(bits_of(w, nbits))])
    idx = np.arange(nbits + 1)
    assert np.array_equal(orc.rank9(D, w, idx), pref.astype(np.uint64))


def py_init_slow(positions, nbits):
    """Independent restatement of select_support_mcl::init_slow content."""
    cap = (nbits + 63) // 64 * 64
    logn = (cap.bit_length() - 1 if cap else 0) + 1
    logn4 = (logn * logn) ** 2 & 0xFFFFFFFF
    m = len(positions)
    sb = (m + 4095) // 4096
    res = {"superblock": [], "mini": [], "long": [], "mini_width": [], "long_width": []}
    for k in range(sb):
        chunk = positions[4096 * k:4096 * k + 4096]
        a, e = chunk[0], chunk[-1]
        res["superblock"].append(a)
        hi = lambda x: (x.bit_length() - 1) if x else 0  # noqa: E731
        if e - a > logn4:
            res["long"].append(list(chunk) + [0] * (4096 - len(chunk)))
            res["mini"].append([0] * 64)
            res["mini_width"].append(0)
            res["long_width"].append(hi(e) + 1)
        else:
            res["long"].append(None)
            mini = [chunk[64 * t] - a if 64 * t < len(chunk) else 0 for t in range(64)]
            res["mini"].append(mini)
            res["mini_width"].append(hi(e - a) + 1)
            res["long_width"].append(0)
    return res, logn, logn4


SEL_CASES = [(1, 0.5, False), (100, 0.5, False), (70000, 0.5, False), (70001, 0.05, False),
             (1 << 20, 0.01, False), (1 << 20, 0.3, True), (300000, 0.002, False)]


@pytest.mark.parametrize("nbits,density,runs", SEL_CASES)
@pytest.mark.parametrize("bit", [0, 1])
def test_select_content_vs_python(orc, nbits, density, runs, bit):
    rng = np.random.default_rng(nbits ^ bit)
    w = rand_words(rng, nbits, density, runs)
    b = bits_of(w, nbits)
    pos = np.nonzero(b == bit)[0].tolist()
    S = orc.Select(w, nbits, bit)
    exp, logn, logn4 = py_init_slow(pos, nbits)
    assert S.arg_cnt == len(pos) and S.logn == logn and S.logn4 == logn4
    assert S.sb == len(exp["superblock"])
    if S.sb == 0:
        return
    assert S.superblock[:S.sb].tolist() == exp["superblock"]
    assert S.mini_width[:S.sb].tolist() == exp["mini_width"]
    assert S.long_width[:S.sb].tolist() == exp["long_width"]
    for k in range(S.sb):
        if exp["long"][k] is None:
            assert S.long_index[k] == -1
            assert S.mini[64 * k:64 * k + 64].tolist() == exp["mini"][k]
        else:
            li = int(S.long_index[k])
            assert li >= 0
            assert S.long_pool[4096 * li:4096 * li + 4096].tolist() == exp["long"][k]


@pytest.mark.parametrize("nbits,density,runs", SEL_CASES)
@pytest.mark.parametrize("bit", [0, 1])
def test_select_query_vs_brute(orc, nbits, density, runs, bit):
    rng = np.random.default_rng(nbits * 3 + bit)
    w = rand_words(rng, nbits, density, runs)
    pos = np.nonzero(bits_of(w, nbits) == bit)[0]
    S = orc.Select(w, nbits, bit)
    if pos.size == 0:
        return
    qs = np.unique(np.concatenate([[1, pos.size], rng.integers(1, pos.size + 1, 3000)]))
    assert np.array_equal(S.queries(qs), pos[qs - 1].astype(np.uint64))
    for i in qs[:50]:
        assert orc.select_brute(w, nbits, bit, int(i)) == pos[i - 1]


def test_long_superblocks_present(orc):
    rng = np.random.default_rng(5)
    w = rand_words(rng, 1 << 20, 0.01)
    S = orc.Select(w, 1 << 20, 1)
    assert S.n_long > 0 and S.n_long <= S.sb


# ---- two-bit patterns `10` / `01` (rank_support_v<t_b,2>, select_support_mcl<t_b,2>)
def pattern_occ(words, nbits_total, pattern):
    """position i > 0 is an occurrence when bits (i-1, i) read the pattern"""
    b = np.unpackbits(words.view(np.uint8), bitorder="little")[:nbits_total].astype(np.int64)
    first, second = (1, 0) if pattern == 10 else (0, 1)
    occ = np.zeros(nbits_total, np.int64)
    occ[1:] = (b[:-1] == first) & (b[1:] == second)
    return occ


def pattern_case(nbits, seed, last_one):
    rng = np.random.default_rng(seed)
    nw = (nbits + 63) // 64
    w = rng.integers(0, np.iinfo(np.uint64).max, nw, dtype=np.uint64, endpoint=True)
    if nbits & 63:
        w[-1] &= np.uint64((1 << (nbits & 63)) - 1)
    if last_one:  # bit nbits-1 set: `10` then has an occurrence in the padding
        w[-1] |= np.uint64(1 << ((nbits - 1) & 63))
    return w


@pytest.mark.parametrize("nbits", [1, 2, 63, 64, 65, 127, 128, 1000, 4099, 100003])
@pytest.mark.parametrize("pattern", [10, 1])
@pytest.mark.parametrize("last_one", [False, True])
def test_pattern_words_vs_bits(orc, nbits, pattern, last_one):
    w = pattern_case(nbits, nbits * 100 + pattern + last_one, last_one)
    cap = w.size * 64
    occ = pattern_occ(w, cap, pattern)  # over whole words (padding bits are 0)
    pw0 = orc.pattern_words(w, nbits, pattern, clip=False)
    pw1 = orc.pattern_words(w, nbits, pattern, clip=True)
    assert np.array_equal(np.unpackbits(pw0.view(np.uint8), bitorder="little")[:cap], occ)
    clipped = occ.copy()
    clipped[nbits:] = 0
    assert np.array_equal(np.unpackbits(pw1.view(np.uint8), bitorder="little")[:cap], clipped)
    if pattern == 10 and last_one and nbits & 63:
        assert occ[nbits] == 1  # the padding occurrence the whole-word count includes
    # rank over whole words: rank(i) = occurrences before i, for every i <= nbits
    D = orc.rank9_build(pw0, nbits)
    pref = np.concatenate([[0], np.cumsum(occ)])
    idx = np.arange(nbits + 1, dtype=np.uint64)
    assert np.array_equal(orc.rank9(D, pw0, idx, nbits), pref[:nbits + 1].astype(np.uint64))
    # select below nbits: the i-th occurrence
    S = orc.Select(pw1, nbits, 1)
    pos = np.nonzero(occ[:nbits])[0]
    assert S.arg_cnt == pos.size
    if pos.size:
        assert np.array_equal(S.queries(np.arange(1, pos.size + 1)), pos.astype(np.uint64))
