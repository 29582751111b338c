"""Pins the oracle's level builder (orc_levels_serial, serialWT restatement).

Checks: (1) the golden fixtures written from the compiled reference builders
(tests/golden/levels_ref.npz, reference wavelet/serialWT/WT.C == levelWT ==
recWT); (2) the compiled reference itself when oracle/_ref is present
(serialWT, levelWT, recWT and, for inputs large enough for it, ddWT);
(3) the independent closed form (stable sort by prefix); (4) the reference's
access-reconstruction check (wavelet/levelWT/WTTime.C:104-113); (5) node
invariants: children partition the parent, child lengths equal the parent's
zero/one counts.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "levels_ref.npz")


def golden_cases():
    z = np.load(GOLD, allow_pickle=False)
    names = sorted({k.split("__")[0] for k in z.files})
    return [(nm, z[nm + "__s"], int(z[nm + "__sigma"][0]), z[nm + "__words"], z[nm + "__reach"],
             z[nm + "__nodes"]) for nm in names]


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_oracle_matches_reference_golden(orc, case):
    name, s, sigma, words, reach, nodes = case
    L = orc.levels_of_sigma(sigma)
    w, nd = orc.levels_serial(s, L)
    assert np.array_equal(w, words)
    r = reach.astype(np.int64)
    assert np.array_equal(nd[r, :5], nodes)
    # exactly the reachable slots are marked reached
    assert set(np.nonzero(nd[:, 5])[0].tolist()) == set(r.tolist())


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c[0])
def test_closed_form_and_access_on_golden(orc, case):
    name, s, sigma, words, reach, nodes = case
    L = orc.levels_of_sigma(sigma)
    assert np.array_equal(orc.levels_closed(s, L), words)
    _, nd = orc.levels_serial(s, L)
    idx = [(int(orc.lib().orc_pbbs_hash(i)) % s.size) for i in range(min(200, s.size))]
    for i in idx:  # wavelet/levelWT/WTTime.C:106-112
        assert orc.wt_query(nd, words, i, sigma) == int(s[i])


REF_CASES = [(100003, 1000, np.uint32, "u32"), (65536, 256, np.uint8, "u8"),
             (1000, 3, np.uint32, "u32"), (250000, 20, np.uint64, "u64"),
             (777, 1, np.uint8, "u8"), (5000, 2, np.uint8, "u8"), (70001, 4099, np.uint64, "u64")]


@pytest.mark.parametrize("n,sigma,dt,sym", REF_CASES)
def test_oracle_matches_compiled_reference(orc, n, sigma, dt, sym):
    if not orc.RefWT("serialWT", sym).ok:
        pytest.skip("oracle/_ref not built (reference sources unavailable)")
    rng = np.random.default_rng(n * 31 + sigma)
    s = rng.integers(0, sigma, n).astype(dt)
    L = orc.levels_of_sigma(sigma)
    w, nd = orc.levels_serial(s, L)
    reached = np.nonzero(nd[:, 5])[0]
    algs = ["serialWT", "levelWT", "recWT"]
    # ddWT divides by zero when n*L < 256*workers (wavelet/ddWT/WT.C:188-189)
    if n * L >= 256 * 64:
        algs.append("ddWT")
    for alg in algs:
        rw, rn = orc.RefWT(alg, sym).build(s, sigma)
        assert np.array_equal(rw, w), alg
        assert np.array_equal(rn[reached], nd[reached, :5]), alg


@pytest.mark.parametrize("seed", range(12))
def test_closed_form_equals_serial_random(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 3000))
    L = int(rng.integers(1, 17))
    dt = [np.uint8, np.uint16, np.uint32, np.uint64][seed % 4]
    hi = min(1 << L, np.iinfo(dt).max + 1)
    s = rng.integers(0, hi, n).astype(dt)
    w, _ = orc.levels_serial(s, L, with_nodes=False)
    assert np.array_equal(w, orc.levels_closed(s, L))


def test_node_invariants(orc):
    rng = np.random.default_rng(7)
    s = rng.integers(0, 300, 20000).astype(np.uint32)
    L = orc.levels_of_sigma(300)
    w, nd = orc.levels_serial(s, L)
    bits = np.unpackbits(w.view(np.uint8), bitorder="little")
    for v in np.nonzero(nd[:, 5])[0]:
        ptr, ln, par, lc, rc = (int(x) for x in nd[v, :5])
        ones = int(bits[ptr:ptr + ln].sum())
        if lc != orc.NONE:
            assert nd[lc, 1] == ln - ones and nd[lc, 2] == v
        else:
            assert ln - ones == 0 or (int(v) + 1).bit_length() == L
        if rc != orc.NONE:
            assert nd[rc, 1] == ones and nd[rc, 2] == v
