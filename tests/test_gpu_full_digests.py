"""Full-size parity of the benchmark inputs against the compiled reference.

The drop-in WT() boundary builds the C1 input (2^24 Zipf bytes, 8 levels) and
both C2 inputs (2^28 uint32 values, uniform and geometric(1e-4), 20 levels) on
the GPU; the level words and the reachable WTnode records must hash to the
SHA-256 digests of the reference levelWT's own output on the same inputs
(wavelet/levelWT/WT.C compiled unmodified; tests/golden/full_size_digests.json,
written by tests/golden/make_digests.py).  The inputs come from the library's
seeded host generators, so they are regenerated here bit for bit."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "full_size_digests.json")))
SEED = 0x5EED2026


def make_input(wsfm, name):
    if name == "c1_zipf_bytes":
        return wsfm.gen_zipf_bytes(SEED + 1, 1 << 24, 1.1)
    if name == "c2_uniform":
        return wsfm.gen_uniform(SEED, 1 << 28, 1 << 20, np.uint32)
    return wsfm.gen_geometric(SEED + 9, 1 << 28, 1e-4, (1 << 20) - 1, np.uint32)


@pytest.mark.parametrize("name", sorted(FIX))
def test_full_size_matches_reference_digest(wsfm, ctx, name):
    spec = FIX[name]
    print("full-size digest", name, spec["n"], spec["levels"], flush=True)
    s = make_input(wsfm, name)
    assert s.size == spec["n"]
    words, nodes = ctx.wt_host(s, spec["sigma"])
    del s
    assert hashlib.sha256(np.ascontiguousarray(words).tobytes()).hexdigest() == spec["words_sha256"]
    del words
    # reachable nodes = the root and every non-empty node (children are set
    # only for non-empty nodes); unreached slots are empty here
    r = np.flatnonzero((nodes[:, 1] > 0) | (np.arange(nodes.shape[0]) == 0)).astype(np.int64)
    assert r.size == spec["nodes_reached"]
    rec = np.ascontiguousarray(np.concatenate([r.astype(np.uint64)[:, None], nodes[r]], axis=1))
    assert hashlib.sha256(rec.tobytes()).hexdigest() == spec["nodes_reached_sha256"]


SUPPORT_CASES = sorted(k for k in FIX if "supports" in FIX[k])


@pytest.mark.parametrize("name", SUPPORT_CASES)
def test_full_size_supports_match_oracle_digest(wsfm, ctx, name):
    """rank_support_v<1,1> + select_support_mcl<1,1> / <0,1> over the whole
    C2 tree (5,368,709,120 bits: positions, superblock starts and mini offsets
    past 2^32), built on the GPU from the device tree, hash to the oracle's
    digests of the reference levelWT's words (rank/rank_support_v_par.hpp:69-123,
    select/select_support_mcl_par.hpp:227-276; wt_int builds them over the
    concatenated tree, wt_int/wt_int_par.hpp:402-404)."""
    spec = FIX[name]
    sup = spec["supports"]
    print("full-size supports", name, sup["nbits"], flush=True)
    s = make_input(wsfm, name)
    n, L = spec["n"], spec["levels"]
    ds = ctx.to_device(s)
    del s
    dw = ctx.empty(wsfm.lib().wtg_levels_words(n, L), np.uint64)
    ctx.levels_build(ds, n, L, dw)
    ds.free()
    assert n * L == sup["nbits"]
    D, s1, s0 = ctx.rank9_select_build(dw, n * L)
    try:
        d = D.get()
        assert d.size == sup["rank9_words"]
        assert hashlib.sha256(d.tobytes()).hexdigest() == sup["rank9_sha256"]
        del d
        for bit, sel in ((1, s1), (0, s0)):
            c = sel.content()
            exp = sup["select%d" % bit]
            assert (c["arg_cnt"], c["sb"], c["n_long"], c["logn"]) == \
                (exp["arg_cnt"], exp["sb"], exp["n_long"], exp["logn"])
            assert wsfm.select_content_sha256(c) == exp["sha256"], "select%d content" % bit
    finally:
        s1.free(); s0.free(); D.free(); dw.free()
