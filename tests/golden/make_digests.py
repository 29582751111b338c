"""Writes tests/golden/full_size_digests.json: SHA-256 digests of the COMPILED
REFERENCE's output for the full-size benchmark inputs.

Run in a container where /root/reference exists and `make -C oracle ref` has
built oracle/_ref (reference wavelet/<alg>/WT.C compiled unmodified, plus
oracle/ref_harness.cpp).  Inputs come from the library's seeded host
generators (splitmix64 streams, the bench's seeds), so the GPU test regenerates
them bit for bit.  For each input the fixture holds the generator call, the
digest of the level words (the reference's ceil(n*L/64) uint64 words) and the
digest of the reachable WTnode records (bitmapPtr, length, parent, leftChild,
rightChild as uint64, UINT_T_MAX widened to 2^64-1, slots in heap order).

For the two C2 inputs (n*L = 5,368,709,120 bits, past 2^32) it also holds the
digests of the supports that sdsl::wt_int builds over the concatenated tree
bits (wt_int/wt_int_par.hpp:402-404): the rank_support_v<1,1> directory
(rank/rank_support_v_par.hpp:69-123, the oracle's restatement orc_rank9_build)
and the select_support_mcl<1,1> / <0,1> content (select_support_mcl_par.hpp:
227-276, the oracle's init_slow restatement), each over the reference
levelWT's own words.  The select digest is select_content_sha256 of the
package (header, superblocks, mini offsets, widths, long index and long pool in
one fixed byte order).

    python tests/golden/make_digests.py             # everything
    python tests/golden/make_digests.py --supports  # only add the support digests
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from oracle import oracle as O  # noqa: E402
import _pkg  # noqa: E402

SEED = 0x5EED2026


def reached(nodes):
    reach, stack = [], [0]
    while stack:
        v = stack.pop()
        reach.append(v)
        for c in (nodes[v, 3], nodes[v, 4]):
            if c != O.NONE:
                stack.append(int(c))
    return np.array(sorted(reach), np.int64)


def digests(words, nodes):
    r = reached(nodes)
    rec = np.ascontiguousarray(np.concatenate([r.astype(np.uint64)[:, None], nodes[r]], axis=1))
    return (hashlib.sha256(np.ascontiguousarray(words).tobytes()).hexdigest(),
            hashlib.sha256(rec.tobytes()).hexdigest(), int(r.size))


def support_digests(pkg, words, nbits):
    """rank9 directory and select1/select0 content digests by the oracle."""
    D = O.rank9_build(words, nbits)
    out = {"nbits": int(nbits), "rank9_words": int(D.size),
           "rank9_sha256": hashlib.sha256(np.ascontiguousarray(D).tobytes()).hexdigest()}
    del D
    for bit in (1, 0):
        S = O.Select(words, nbits, bit)
        c = {"arg_cnt": S.arg_cnt, "sb": S.sb, "n_long": S.n_long, "logn": S.logn, "logn4": S.logn4,
             "superblock": S.superblock, "mini": S.mini, "mini_width": S.mini_width,
             "long_width": S.long_width, "long_index": S.long_index, "long_pool": S.long_pool}
        out["select%d" % bit] = {"arg_cnt": S.arg_cnt, "sb": S.sb, "n_long": S.n_long, "logn": S.logn,
                                 "sha256": pkg.select_content_sha256(c)}
        del S, c
    return out


def main():
    pkg = _pkg.load()
    only_supports = "--supports" in sys.argv[1:]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "full_size_digests.json")
    n2, n1 = 1 << 28, 1 << 24
    cases = [
        ("c1_zipf_bytes", "gen_zipf_bytes(seed=0x5EED2027, n=2^24, s=1.1)", 256, "u8",
         lambda: pkg.gen_zipf_bytes(SEED + 1, n1, 1.1)),
        ("c2_uniform", "gen_uniform(seed=0x5EED2026, n=2^28, range=2^20, uint32)", 1 << 20, "u64",
         lambda: pkg.gen_uniform(SEED, n2, 1 << 20, np.uint32)),
        ("c2_geometric", "gen_geometric(seed=0x5EED202F, n=2^28, p=1e-4, cap=2^20-1, uint32)", 1 << 20, "u64",
         lambda: pkg.gen_geometric(SEED + 9, n2, 1e-4, (1 << 20) - 1, np.uint32)),
    ]
    out = json.load(open(path)) if only_supports else {}
    for name, gen, sigma, sym, make in cases:
        if only_supports and not name.startswith("c2_"):
            continue
        s = make()
        ref = O.RefWT("levelWT", sym)
        assert ref.ok, "build oracle/_ref first (make -C oracle ref)"
        words, nodes = ref.build(s.astype(np.uint64) if sym == "u64" else s, sigma)
        wd, nd, nr = digests(words, nodes)
        if only_supports:
            assert out[name]["words_sha256"] == wd
        else:
            out[name] = {"generator": gen, "n": int(s.size), "sigma": sigma,
                         "levels": O.levels_of_sigma(sigma), "reference": "wavelet/levelWT/WT.C (%s)" % sym,
                         "words_sha256": wd, "nodes_reached_sha256": nd, "nodes_reached": nr}
        print(name, wd, nr, flush=True)
        if name.startswith("c2_"):
            out[name]["supports"] = support_digests(pkg, words, int(s.size) * O.levels_of_sigma(sigma))
            print(name, out[name]["supports"], flush=True)
        del s, words, nodes
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
