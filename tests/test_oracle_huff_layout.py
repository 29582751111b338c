"""The wt_huff tree shape and bit layout of the oracle (orc_huff_shape), pinned
by what the reference states:

* the Huffman merge of wt_huff/wt_huff_par.hpp:92-117: a min priority queue of
  (frequency, node id) pairs (std::pair order: ties on frequency go to the
  smaller id), leaves pushed in symbol order with ids 0..s-1, two pops per
  merge (first pop = child[0]), the parent taking the next id.  Checked against
  an independent heapq restatement, on tie-heavy alphabets where a wrong tie
  rule changes the tree;
* "Convert code tree into BFS order in memory and calculate bv_pos values"
  (wt_huff/wt_pc_par.hpp:134-135): node ids are breadth-first from the root
  (root 0, depth non-decreasing with the id, the children of a node before
  those of any later node, child[0] before child[1]); the inner nodes' runs
  are contiguous in that order, bv_pos[v] = total frequency of the inner nodes
  before v, and the bit vector holds sum_c C[c] * |code(c)| bits;
* the bits of build_recursive (wt_pc_par.hpp:193-257): the bit of an
  occurrence at a node of depth l is the l-th branch of its root-to-leaf path,
  and each node's occurrences stay in text order.

The byte order sdsl's own byte_tree uses beyond this statement stays parity
unpinned (sdsl-lite is not available)."""
import heapq
from collections import deque

import numpy as np
import pytest


def heap_shape(C):
    """Independent restatement: (children, symbols, freq) in BFS ids."""
    freq, kids, sym = [], [], []
    pq = []
    for c in range(256):
        if C[c] > 0:
            heapq.heappush(pq, (int(C[c]), len(freq)))
            freq.append(int(C[c])); kids.append(None); sym.append(c)
    while len(pq) > 1:
        f1, v1 = heapq.heappop(pq)
        f2, v2 = heapq.heappop(pq)
        heapq.heappush(pq, (f1 + f2, len(freq)))
        freq.append(f1 + f2); kids.append((v1, v2)); sym.append(-1)
    if not freq:
        return [], [], []
    order, q = [], deque([len(freq) - 1])
    while q:
        v = q.popleft()
        order.append(v)
        if kids[v] is not None:
            q.extend(kids[v])
    new = {old: i for i, old in enumerate(order)}
    ch = [(new[kids[o][0]], new[kids[o][1]]) if kids[o] is not None else (-1, -1) for o in order]
    return ch, [sym[o] for o in order], [freq[o] for o in order]


def alphabets():
    rng = np.random.default_rng(2026)
    cases = {
        "equal16": np.repeat(np.arange(16), 5),
        "equal256_ones": np.arange(256),
        "powers_of_two": np.concatenate([np.full(1 << k, 40 + k) for k in range(9)]),
        "fibonacci": np.concatenate([np.full(f, i) for i, f in enumerate([1, 1, 2, 3, 5, 8, 13, 21, 34, 55])]),
        "pairs_tied": np.concatenate([np.full(3, c) for c in range(0, 200, 7)] +
                                     [np.full(6, c) for c in range(3, 200, 11)]),
        "dna_like": np.array([0] + [65] * 30 + [67] * 20 + [71] * 20 + [84] * 30 + [78] * 1),
        "zipf": rng.zipf(1.3, 5000).clip(0, 255),
        "two": np.array([7, 9, 9]),
        "one": np.array([42] * 10),
    }
    return {k: v.astype(np.uint8) for k, v in cases.items()}


CASES = alphabets()


@pytest.mark.parametrize("name", sorted(CASES))
def test_shape_equals_independent_heap(orc, name):
    t = CASES[name]
    C = np.bincount(t, minlength=256)
    H = orc.Huff(t).H
    ch, sym, freq = heap_shape(C)
    assert H.n_nodes == len(ch)
    for v in range(H.n_nodes):
        assert (H.child0[v], H.child1[v]) == ch[v], v
        assert H.sym[v] == sym[v], v
        assert H.freq[v] == freq[v], v


@pytest.mark.parametrize("name", sorted(CASES))
def test_bfs_order_and_bv_pos(orc, name):
    t = CASES[name]
    C = np.bincount(t, minlength=256)
    H = orc.Huff(t).H
    n = H.n_nodes
    depth = [0] * n
    assert H.parent[0] == -1
    nxt = 1
    for v in range(n):
        if H.child0[v] >= 0:
            # children come next in BFS order, child[0] first
            assert (H.child0[v], H.child1[v]) == (nxt, nxt + 1)
            nxt += 2
            for c in (H.child0[v], H.child1[v]):
                assert H.parent[c] == v
                depth[c] = depth[v] + 1
        else:
            assert H.child1[v] == -1 and H.sym[v] >= 0 and H.c_to_leaf[H.sym[v]] == v
    assert nxt == n
    assert all(depth[v] <= depth[v + 1] for v in range(n - 1))
    acc = 0
    for v in range(n):  # inner runs contiguous in BFS order
        if H.child0[v] >= 0:
            assert H.bv_pos[v] == acc
            assert H.freq[v] == H.freq[H.child0[v]] + H.freq[H.child1[v]]
            acc += H.freq[v]
    lens = np.array([H.path_len[c] for c in range(256)])
    assert H.bv_size == acc == int((C * lens).sum())
    for c in np.flatnonzero(C):
        v, path = H.c_to_leaf[c], []
        while H.parent[v] >= 0:
            p = H.parent[v]
            path.append(1 if H.child1[p] == v else 0)
            v = p
        path = path[::-1]
        assert H.path_len[c] == len(path) == depth[H.c_to_leaf[c]]
        assert all(((H.path[c] >> l) & 1) == b for l, b in enumerate(path))


@pytest.mark.parametrize("name", ["dna_like", "pairs_tied", "zipf", "equal16"])
def test_bits_follow_paths_in_text_order(orc, name):
    rng = np.random.default_rng(7)
    t = rng.permutation(CASES[name])
    Hf = orc.Huff(t)
    H = Hf.H
    bits = np.unpackbits(Hf.words.view(np.uint8), bitorder="little")[:Hf.nbits]
    fill = {}
    for x in t.tolist():
        v = 0
        for l in range(H.path_len[x]):
            b = (H.path[x] >> l) & 1
            k = fill.get(v, 0)
            assert bits[H.bv_pos[v] + k] == b
            fill[v] = k + 1
            v = H.child1[v] if b else H.child0[v]
    for v, k in fill.items():
        assert k == H.freq[v]
