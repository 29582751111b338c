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
