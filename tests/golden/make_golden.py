"""Writes tests/golden/levels_ref.npz from the COMPILED REFERENCE builders.

Run in a container where /root/reference exists and `make -C oracle ref` has
built oracle/_ref/libref_<alg>_<sym>.so (reference wavelet/<alg>/WT.C compiled
unmodified, plus oracle/ref_harness.cpp).  For each case the fixture holds the
input symbols (seeded numpy), the reference serialWT words, and the reachable
WTnode records (bitmapPtr, length, parent, leftChild, rightChild; UINT_T_MAX
widened to 2^64-1).  levelWT and recWT are checked identical before writing.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

CASES = [
    # name, n, sigma, dtype, sym, generator
    ("u8_zipf", 3000, 256, np.uint8, "u8", "zipf"),
    ("u8_sigma1", 777, 1, np.uint8, "u8", "uniform"),
    ("u8_sigma2", 5000, 2, np.uint8, "u8", "uniform"),
    ("u8_sigma3_odd", 4097, 3, np.uint8, "u8", "uniform"),
    ("u8_sigma200", 12345, 200, np.uint8, "u8", "uniform"),
    ("u32_sigma1000", 8011, 1000, np.uint32, "u32", "uniform"),
    ("u32_sigma3", 1000, 3, np.uint32, "u32", "uniform"),
    ("u32_sigma2p20", 2500, 1 << 20, np.uint32, "u32", "uniform"),
    ("u32_sigma5_tiny", 7, 5, np.uint32, "u32", "uniform"),
    ("u64_sigma20", 25000, 20, np.uint64, "u64", "uniform"),
    ("u64_sigma70000_geom", 8000, 70000, np.uint64, "u64", "geom"),
]


def make_input(rng, n, sigma, dtype, kind):
    if kind == "zipf":
        r = rng.zipf(1.3, n)
        return ((r - 1) % sigma).astype(dtype)
    if kind == "geom":
        return np.minimum(rng.geometric(1e-4, n) - 1, sigma - 1).astype(dtype)
    return rng.integers(0, sigma, n).astype(dtype)


def main():
    rng = np.random.default_rng(0x5EED2026)
    out = {}
    for name, n, sigma, dt, sym, kind in CASES:
        s = make_input(rng, n, sigma, dt, kind)
        ref = O.RefWT("serialWT", sym)
        assert ref.ok, "build oracle/_ref first (make -C oracle ref)"
        words, nodes = ref.build(s, sigma)
        for alg in ("levelWT", "recWT"):
            w2, n2 = O.RefWT(alg, sym).build(s, sigma)
            assert (w2 == words).all(), (alg, name)
        # reachable nodes: serialWT leaves unreached slots uninitialised, so
        # take the reachable set from a walk over the child pointers.
        reach, stack = [], [0]
        while stack:
            v = stack.pop()
            reach.append(v)
            for c in (nodes[v, 3], nodes[v, 4]):
                if c != O.NONE:
                    stack.append(int(c))
        reach = np.array(sorted(reach), np.uint64)
        out[name + "__s"] = s
        out[name + "__sigma"] = np.array([sigma], np.uint64)
        out[name + "__words"] = words
        out[name + "__reach"] = reach
        out[name + "__nodes"] = nodes[reach.astype(np.int64)]
        print(name, n, sigma, words.size, reach.size)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "levels_ref.npz"),
                        **out)


if __name__ == "__main__":
    main()
