"""Full-size outputs of the benchmark workloads against committed digests of the
reference's (or brute force's) output on the same inputs
(tests/golden/workload_digests.json, written by make_workload_digests.py):

  C1  10^6 wt_int rank(c, i) answers over the 2^24 Zipf bytes
  C2  10^7 rank1 and 10^7 select1 answers over the 5.37-Gbit C2 tree
  C3  the GPU suffix array and BWT of the 2^27-byte DNA text against the
      reference's DC3 suffix array (suffixArray/parallelKS/pks.C), and the
      10^6 FM-index counts against binary search in that suffix array
  C4  every sequence of the 4096-sequence batch against the reference levelWT
  C5  the 10^8 learned-select answers (and the plain select's) against brute force

Inputs and queries come from the package's workloads module (seeded)."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
WD = json.load(open(os.path.join(HERE, "golden", "workload_digests.json")))
FIX = json.load(open(os.path.join(HERE, "golden", "full_size_digests.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_c1_rank_answers(wsfm, ctx):
    W = wsfm.workloads
    print("C1 rank answers", flush=True)
    s = W.c1_input()
    d = ctx.to_device(s)
    wt = ctx.wtint_build(d, W.C1_N, W.C1_LEVELS)
    qi, qc = W.c1_queries()
    try:
        got = wt.rank(qi, qc)
        assert sha(got) == WD["c1"]["rank_answers_sha256"]
        assert sha(wt.words()[:wsfm.lib().wtg_levels_words(W.C1_N, W.C1_LEVELS)]) == \
            FIX["c1_zipf_bytes"]["words_sha256"]
    finally:
        wt.free(); d.free()


def test_c2_rank_select_answers(wsfm, ctx):
    W = wsfm.workloads
    spec = WD["c2"]
    print("C2 rank1/select1 answers", spec["nbits"], flush=True)
    s = W.c2_input("uniform")
    ds = ctx.to_device(s)
    del s
    dw = ctx.empty(wsfm.lib().wtg_levels_words(W.C2_N, W.C2_LEVELS), np.uint64)
    ctx.levels_build(ds, W.C2_N, W.C2_LEVELS, dw)
    ds.free()
    nbits = W.C2_N * W.C2_LEVELS
    D, s1, s0 = ctx.rank9_select_build(dw, nbits)
    try:
        assert s1.arg_cnt == spec["ones"]
        qi, qj = W.c2_queries(nbits, s1.arg_cnt)
        di, dj = ctx.to_device(qi), ctx.to_device(qj)
        r = ctx.rank9_batch(dw, D, nbits, di)
        sl = ctx.select_batch(s1, dw, dj)
        assert sha(r.get()) == spec["rank1_answers_sha256"]
        assert sha(sl.get()) == spec["select1_answers_sha256"]
        for a in (di, dj, r, sl):
            a.free()
    finally:
        s1.free(); s0.free(); D.free(); dw.free()


def test_c3_suffix_array_bwt_counts(wsfm, ctx):
    W = wsfm.workloads
    spec = WD["c3"]
    print("C3 GPU suffix array, BWT, FM counts", spec["n"], flush=True)
    t = W.c3_text()
    pats = W.c3_patterns(t)
    dt = ctx.to_device(t)
    dsa = wsfm.suffix_array_gpu(ctx, dt, t.size)
    dbwt = wsfm.bwt_from_sa_gpu(ctx, dt, dsa, t.size)
    try:
        assert sha(dsa.get()) == spec["sa_sha256"]
        assert sha(dbwt.get()) == spec["bwt_sha256"]
        wt = ctx.huff_build(dbwt)
        try:
            got = wt.fm_count(pats)
            assert int(got.sum()) == spec["counts_sum"]
            assert sha(got.astype(np.uint64)) == spec["counts_sha256"]
        finally:
            wt.free()
    finally:
        dsa.free(); dbwt.free(); dt.free()


def test_c4_batch_words(wsfm, ctx):
    W = wsfm.workloads
    spec = WD["c4"]
    print("C4 batch words", spec["seqs"], spec["symbols"], flush=True)
    n, sigma, sym, off = W.c4_batch()
    assert int(n.sum()) == spec["symbols"]
    d = ctx.to_device(sym)
    del sym
    B = wsfm.Batch(ctx, d, n, sigma, off)
    try:
        B.run()
        h = hashlib.sha256()
        for i in range(n.size):
            h.update(B.extract(i).tobytes())
        assert h.hexdigest() == spec["words_sha256"]
    finally:
        B.free(); d.free()


def test_c5_learned_select_answers(wsfm, ctx):
    W = wsfm.workloads
    spec = WD["c5"]
    print("C5 learned select answers", spec["queries"], flush=True)
    w, _ = W.c5_vectors()
    nvec, nw = w.shape
    dv = []
    for b in range(nvec):
        p = np.zeros(nw + 1, np.uint64)
        p[:nw] = w[b]
        dv.append(ctx.to_device(p))
    del w
    ls = wsfm.LearnedSelect(ctx, dv, [nw * 64] * nvec)
    try:
        ones = [ls.info(b)["ones"] for b in range(nvec)]
        assert ones == spec["ones"]
        vec, j = W.c5_queries(np.array(ones, np.uint64))
        dvq, djq = ctx.to_device(vec), ctx.to_device(j)
        del vec, j
        out = ctx.empty(djq.size, np.uint64)
        for learned in (True, False):
            ls.query(dvq, djq, out, learned)
            assert sha(out.get()) == spec["answers_sha256"], "learned" if learned else "plain"
        for a in (dvq, djq, out):
            a.free()
    finally:
        ls.free()
        for a in dv:
            a.free()
