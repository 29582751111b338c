#!/usr/bin/env python3
"""Benchmark of the wavelet-tree / rank / select path on one MI355X.

Default run (one JSON line): the headline workload, the integer wavelet tree of
sdsl::wt_int<> with max_level = 20 over n = 2^28 uint32 symbols uniform in
[0, 2^20) (seed 0x5EED2026).  One step = one full tree build (the 20 level
passes) with the input already resident in HBM.  `value` is Gsym/s = n / step
time; `roofline` is the dominant kernel's (the level split: algorithmic bytes
2*n*w + n/8 per launch over its HIP-event-timed average duration);
`build_roofline` is the whole build against B_a = n*(2*L*w + L/8).  The line
also carries, each checked against committed SHA-256 digests of the reference's
(or brute force's) output on the same inputs (tests/golden/*.json):

  parity          the timed words, both C2 inputs, their rank9 / select1 /
                  select0 supports and 10^7 rank1 + 10^7 select1 answers
  wt_int_build    the full construction (tree + rank9 + select1 + select0)
  learned_select  64 bitvectors x 2^26 bits, 10^8 queries: learned vs plain
                  sampled select (and the placed-only model), sectors per query
  c1_bytes        levelWT bytes, n = 2^24 Zipf, 8 levels + rank9, 10^6 rank(c,i)
  c3_fm           GPU suffix array + BWT + wt_huff of 2^27 DNA, 10^6 FM counts
  c4_batch        4096 uneven byte sequences in one hipGraph replay

cpu_baseline: the reference levelWT and recWT (wavelet/{levelWT,recWT}/WT.C,
compiled with their OpenMP backend and -DINT -DLONG into oracle/_ref) timed on
this host's cores on a bounded prefix of the same input; the faster one is the
value.  `--workload wt_int|bytes|fm|batch|lsel` prints one workload's line only.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "WT build Gsym/s & %HBM peak; select Mq/s learned vs plain (bit-exact)"
SEED = 0x5EED2026
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_pkg():
    import importlib.util
    d = os.path.join(ROOT, "wavelet-suffix-fm-index_amd")
    name = "wavelet_suffix_fm_index_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    m = importlib.util.module_from_spec(spec)
    sys.modules[name] = m
    spec.loader.exec_module(m)
    return m


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def golden(name):
    p = os.path.join(GOLDEN, name)
    return json.load(open(p)) if os.path.exists(p) else {}


def sha(a, dt=None):
    a = np.ascontiguousarray(a if dt is None else a.astype(dt, copy=False))
    return hashlib.sha256(a.tobytes()).hexdigest()


def frac(bytes_, ms):
    return round(bytes_ / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)


def timed(ctx, fn, steps):
    ctx.timer_start()
    for _ in range(steps):
        fn()
    return ctx.timer_stop() / steps


# ---------------------------------------------------------------- CPU side
def cpu_baseline_levels(s, L, sample_n, rounds, sym):
    """The reference levelWT and recWT (oracle/_ref, OpenMP) on the first
    sample_n symbols; the oracle's serial restatement when they are not built."""
    from oracle import oracle
    sample = np.ascontiguousarray(s[:sample_n].astype(np.uint64) if sym == "u64" else s[:sample_n])
    res = {}
    for alg in ("levelWT", "recWT"):
        ref = oracle.RefWT(alg, sym)
        if ref.ok:
            res[alg] = (sample_n / float(ref.time(sample, 1 << L, rounds=rounds).min()) / 1e9, ref.workers())
    if res:
        best = max(res, key=lambda k: res[k][0])
        return {"value": round(res[best][0], 5), "unit": "Gsym/s", "cores": int(res[best][1]),
                "kind": "reference", "best": best,
                "levelWT_Gsym_s": round(res["levelWT"][0], 5) if "levelWT" in res else None,
                "recWT_Gsym_s": round(res["recWT"][0], 5) if "recWT" in res else None,
                "sample": "reference levelWT and recWT (OpenMP%s) on the first %d symbols of the same "
                          "input, L=%d, best of %d rounds each; value = the faster"
                          % (", -DINT -DLONG" if sym == "u64" else ", byte symbols", sample_n, L, rounds)}
    t0 = time.perf_counter()
    oracle.levels_serial(sample, L, with_nodes=False)
    return {"value": round(sample_n / (time.perf_counter() - t0) / 1e9, 5), "unit": "Gsym/s", "cores": 1,
            "kind": "port", "sample": "serialWT restatement on the first %d symbols" % sample_n}


def measured_traffic(log2n, L):
    """HBM bytes per multi-level emit launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by a 1 GiB copy), for the
    workload it was measured on; None otherwise."""
    path = os.path.join(ROOT, "profiles", "mld_emit_u32_n268M_L20.md")
    if (log2n, L) != (28, 20) or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch(mld_emit) = "):
            return float(line.split("=")[1])
    return None


# --------------------------------------------------------------- C2 supports
def supports_and_queries(pkg, ctx, dw, nbits, nq, fix_sup, wd_c2):
    """rank_support_v + select_support_mcl (1 and 0) over all level bits
    (timed), digests of the three structures, then nq uniform rank1(i) and nq
    uniform select1(j) queries (device batches), their answer digests."""
    W = pkg.workloads
    ctx.set_blocking(True)
    for _ in range(2):  # warm
        D, s1, s0 = ctx.rank9_select_build(dw, nbits)
        s1.free(); s0.free(); D.free()
    ctx.timer_start()
    D, s1, s0 = ctx.rank9_select_build(dw, nbits)
    build_ms = ctx.timer_stop()
    par = {}
    if fix_sup:
        par["rank9_digest_match"] = sha(D.get()) == fix_sup["rank9_sha256"]
        par["select1_digest_match"] = pkg.select_content_sha256(s1.content()) == fix_sup["select1"]["sha256"]
        par["select0_digest_match"] = pkg.select_content_sha256(s0.content()) == fix_sup["select0"]["sha256"]
    res = {"rank9_select_build_ms": round(build_ms, 3), "nbits": nbits}
    if nq:
        qi, qj = W.c2_queries(nbits, s1.arg_cnt, nq)
        di, dj = ctx.to_device(qi), ctx.to_device(qj)
        del qi, qj
        out_r, out_s = ctx.empty(nq, np.uint64), ctx.empty(nq, np.uint64)
        for name, fn in (("rank", lambda: ctx.rank9_batch(dw, D, nbits, di, out_r)),
                         ("select1", lambda: ctx.select_batch(s1, dw, dj, out_s))):
            fn()
            t = timed(ctx, fn, 3)
            res[name + "_Mq_s"] = round(nq / (t / 1e3) / 1e6, 1)
        res["queries"] = nq
        if wd_c2 and nq == W.C2_QUERIES:
            par["rank1_answers_digest_match"] = sha(out_r.get()) == wd_c2["rank1_answers_sha256"]
            par["select1_answers_digest_match"] = sha(out_s.get()) == wd_c2["select1_answers_sha256"]
        for a in (di, dj, out_r, out_s):
            a.free()
    s1.free(); s0.free(); D.free()
    return res, par


def bench_wt_int(pkg, args, full):
    W = pkg.workloads
    n, L, w = 1 << args.log2n, args.levels, 4
    fix = golden("full_size_digests.json")
    wd = golden("workload_digests.json")
    full_size = (n, L) == (W.C2_N, W.C2_LEVELS)
    log("generating n=%d uint32 symbols in [0,2^%d)" % (n, L))
    s = W.c2_input("uniform") if full_size else pkg.gen_uniform(SEED, n, 1 << L, np.uint32)
    ctx = pkg.Context(0)
    if args.calibrate_copy:
        a, b = ctx.empty(1 << 27, np.uint64), ctx.empty(1 << 27, np.uint64)
        a.zero()
        for _ in range(2):
            pkg._check(pkg.lib().wtg_memcpy_d2d(ctx.h, b.ptr, a.ptr, 1 << 30))
        ctx.sync()
        a.free(); b.free()
    d = ctx.to_device(s)
    nw = pkg.lib().wtg_levels_words(n, L)
    dw = ctx.empty(nw, np.uint64)
    ctx.set_blocking(False)
    log("warmup %d" % args.warmup)
    for _ in range(args.warmup):
        ctx.levels_build(d, n, L, dw)
    ctx.sync()
    log("timing %d steps" % args.steps)
    t0 = time.perf_counter()
    ctx.timer_start()
    for _ in range(args.steps):
        ctx.levels_build(d, n, L, dw)
    ms = ctx.timer_stop()
    ctx.sync()
    wall = time.perf_counter() - t0
    # per-kernel times in separate builds (events around the multi-level kernels only)
    ksteps = 5
    ctx.kernel_timing_reset()
    ctx.kernel_timing_only("mld_")
    ctx.kernel_timing(True)
    for _ in range(ksteps):
        ctx.levels_build(d, n, L, dw)
    kt = {k: ctx.kernel_time("mld_" + k) for k in ("count", "scan", "hist", "cum", "emit", "ntiles", "tiles")}
    ctx.kernel_timing(False)
    ctx.kernel_timing_only(None)
    # the level-by-level split pipeline on the same input, for comparison
    ctx.set_level_method(1)
    ctx.levels_build(d, n, L, dw)
    split_ms_step = timed(ctx, lambda: ctx.levels_build(d, n, L, dw), 3)
    ctx.set_level_method(0)
    ctx.levels_build(d, n, L, dw)
    ctx.sync()
    ctx.set_blocking(True)

    parity = {}
    words = dw.get()
    if full_size and "c2_uniform" in fix:
        parity["words_digest_match"] = sha(words) == fix["c2_uniform"]["words_sha256"]
    else:  # size-independent check: ones per level
        bits = np.unpackbits(words.view(np.uint8), bitorder="little")
        parity["level_popcounts_match"] = all(
            int(bits[l * n:(l + 1) * n].sum()) == int(((s >> (L - 1 - l)) & 1).sum()) for l in range(L))
        del bits
    del words
    log("supports and queries")
    sup, spar = supports_and_queries(pkg, ctx, dw, n * L, args.queries,
                                     fix.get("c2_uniform", {}).get("supports") if full_size else None,
                                     wd.get("c2") if full_size else None)
    parity.update(spar)

    # the full sdsl::wt_int construction: tree + rank9 + select1 + select0 from
    # the device symbols (wt_int/wt_int_par.hpp:368-406)
    wt = ctx.wtint_build(d, n, L)
    wt.free()
    ctx.timer_start()
    wt = ctx.wtint_build(d, n, L)
    wt_ms = ctx.timer_stop()
    wt.free()
    nbits = n * L
    B_a = n * (2 * L * w + L / 8.0)
    B_sup = nbits / 8.0 + nbits / 32.0 + 2 * nbits / 8.0 + 2 * (nbits / 64.0)  # rank9 + two select reads + minis
    wt_int = {"ms": round(wt_ms, 4), "Gsym_s": round(n / (wt_ms / 1e3) / 1e9, 4),
              "bytes_model": B_a + B_sup, "frac": frac(B_a + B_sup, wt_ms),
              "note": "levels + rank_support_v + select_support_mcl<1> + <0>; bytes = B_a + N/8 + N/32 "
                      "(rank9) + 2*(N/8 + ~N/64 samples) (select1, select0)"}

    geo = None
    if args.geometric:
        log("geometric input")
        g = W.c2_input("geometric") if full_size else pkg.gen_geometric(SEED + 9, n, 1e-4, (1 << L) - 1, np.uint32)
        dg = ctx.to_device(g)
        ctx.levels_build(dg, n, L, dw)
        ctx.set_blocking(False)
        gms = timed(ctx, lambda: ctx.levels_build(dg, n, L, dw), 3)
        ctx.sync()
        ctx.set_blocking(True)
        geo = {"Gsym_s": round(n / (gms / 1e3) / 1e9, 4), "ms": round(gms, 4)}
        if full_size and "c2_geometric" in fix:
            geo["words_digest_match"] = sha(dw.get()) == fix["c2_geometric"]["words_sha256"]
            _, gpar = supports_and_queries(pkg, ctx, dw, n * L, 0, fix["c2_geometric"].get("supports"), None)
            geo.update(gpar)
        dg.free()
        del g
    d.free()
    dw.free()
    ctx.close()

    step_ms = ms / args.steps
    # the multi-level direct passes: groups of up to 8 levels; per group the
    # count kernel reads the keys, the emit kernel reads them, writes the group's
    # level bits and (all but the last group) the keys in the next order (uint32)
    groups, s_ = [], 0
    while s_ < L:
        g_ = min(8, L - s_)
        groups.append(g_)
        s_ += g_
    kw = [w] + [4] * (len(groups) - 1)
    emit_bytes = sum(n * kw[i] + (4 * n if i + 1 < len(groups) else 0) + g_ * n / 8.0
                     for i, g_ in enumerate(groups))
    count_bytes = sum(n * kw[i] for i in range(len(groups)))
    emit_ms, emit_n = kt["emit"]
    emit_s_step = emit_ms / ksteps / 1e3
    achieved = emit_bytes / emit_s_step / 1e9
    kms = {k: round(v[0] / ksteps, 4) for k, v in kt.items()}
    out = {
        "metric": METRIC, "value": round(n / (step_ms / 1e3) / 1e9, 4), "unit": "Gsym/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(step_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "wt_int tree build (levelWT levels), n=2^%d uint32 uniform "
                               "in [0,2^%d), max_level=%d" % (args.log2n, L, L),
                   "n": n, "levels": L, "sym_bytes": w, "seed": SEED},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(args.log2n, L),
                     "kernel": "k_mld_emit (multi-level pass: %s levels per launch)" % "+".join(map(str, groups)),
                     "bytes_per_launch": round(emit_bytes / len(groups), 1),
                     "avg_launch_ms": round(emit_ms / max(emit_n, 1), 4), "launches": emit_n},
        "build_roofline": {"B_a_bytes": B_a, "frac": frac(B_a, step_ms),
                           "bytes_moved_model": emit_bytes + count_bytes,
                           "frac_of_moved": frac(emit_bytes + count_bytes, step_ms),
                           "note": "B_a = n*(2*L*w + L/8) is the level-by-level model; the multi-level passes "
                                   "move bytes_moved_model (keys read twice and written once per group of up "
                                   "to 8 levels, plus the bits)",
                           "kernels_ms_per_step": kms},
        "level_split_pipeline": {"ms": round(split_ms_step, 4),
                                 "Gsym_s": round(n / (split_ms_step / 1e3) / 1e9, 4),
                                 "frac_B_a": frac(B_a, split_ms_step)},
        "parity": parity,
        "wt_int_build": wt_int,
        "rank_select": sup,
        "geometric_input": geo,
        "wall_s": round(wall, 3),
    }
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline_levels(s, L, 1 << args.cpu_log2n, 3, "u64")
    del s
    ok = all(v for v in parity.values()) and (geo is None or all(v for k, v in geo.items() if k.endswith("match")))
    return out, ok


# --------------------------------------------------------------- C5
def bench_lsel(pkg, args, full):
    """Learned vs plain sampled select: 64 bitvectors of 2^26 bits, uniform
    (vector, rank) queries; both answer sets must be identical (and equal the
    committed brute-force digest at the full query count)."""
    W = pkg.workloads
    nvec, nb, nq = W.C5_VECS, 1 << args.lsel_log2bits, args.lsel_queries
    log("generating %d bitvectors of %d bits" % (nvec, nb))
    w, dens = W.c5_vectors(nvec, args.lsel_log2bits)
    ctx = pkg.Context(0)
    nw = w.shape[1]
    dv = []
    for b in range(nvec):
        padded = np.zeros(nw + 1, np.uint64)
        padded[:nw] = w[b]
        dv.append(ctx.to_device(padded))
    del w
    log("build + train")
    prm = pkg.LselParams(16, args.lsel_gap, args.lsel_tol, args.lsel_epochs, 4.0, 512)
    # the first build in the process also loads the kernels and maps fresh
    # memory (reported as cold); the one kept is the second
    ctx.sync()
    t0 = time.perf_counter()
    pkg.LearnedSelect(ctx, dv, [nb] * nvec, prm).free()
    build_cold_s = time.perf_counter() - t0
    ctx.sync()
    t0 = time.perf_counter()
    ls = pkg.LearnedSelect(ctx, dv, [nb] * nvec, prm)
    build_s = time.perf_counter() - t0
    infos = [ls.info(b) for b in range(nvec)]
    ones = np.array([i["ones"] for i in infos], np.uint64)
    vec, j = W.c5_queries(ones, nq)
    dvq, djq = ctx.to_device(vec), ctx.to_device(j)
    del vec, j
    out_p, out_l = ctx.empty(nq, np.uint64), ctx.empty(nq, np.uint64)
    res = {}
    ctx.kernel_timing_reset()
    for name, learned, o in (("plain", False, out_p), ("learned", True, out_l)):
        ls.query(dvq, djq, o, learned)
        ms = timed(ctx, lambda: ls.query(dvq, djq, o, learned), args.steps)
        res[name] = {"Mq_s": round(nq / (ms / 1e3) / 1e6, 1), "ms": round(ms, 3)}
    same = bool(np.array_equal(out_p.get(), out_l.get()))
    wd = golden("workload_digests.json").get("c5")
    digest_ok = None
    if wd and nq == W.C5_QUERIES and args.lsel_log2bits == W.C5_LOG2BITS:
        digest_ok = sha(out_l.get()) == wd["answers_sha256"]
    # sectors of HBM per query on each path, over the first min(nq, 10^7) queries
    ns = min(nq, 10_000_000)
    sec = {}
    for name, learned, o in (("plain", False, out_p), ("learned", True, out_l)):
        dsec = ls.sectors(dvq, djq, o, learned)
        a = dsec.get()[:ns]
        dsec.free()
        sec[name] = round(float(a.mean()), 3)
    # the placed-only model (no training) on the same vectors and queries
    placed = None
    if args.lsel_epochs:
        prm0 = pkg.LselParams(16, args.lsel_gap, args.lsel_tol, 0, 4.0, 512)
        ctx.sync()
        t0 = time.perf_counter()
        ls0 = pkg.LearnedSelect(ctx, dv, [nb] * nvec, prm0)
        build0_s = time.perf_counter() - t0
        out0 = ctx.empty(nq, np.uint64)
        ls0.query(dvq, djq, out0, True)
        ms0 = timed(ctx, lambda: ls0.query(dvq, djq, out0, True), args.steps)
        dsec = ls0.sectors(dvq, djq, out0, True)
        a = dsec.get()[:ns]
        dsec.free()
        placed = {"Mq_s": round(nq / (ms0 / 1e3) / 1e6, 1), "sectors_per_query": round(float(a.mean()), 3),
                  "build_s": round(build0_s, 4),
                  "identical_answers": bool(np.array_equal(out0.get(), out_p.get()))}
        i0 = [ls0.info(b) for b in range(nvec)]
        placed["model_bytes"] = sum(i["nknots"] * 8 + ((i["ones"] >> i["lut_log"]) + 1) * 4 for i in i0 if i["learned"])
        out0.free()
        ls0.free()
    nl = sum(1 for i in infos if i["learned"])
    model_bytes = sum(i["nknots"] * 8 + ((i["ones"] >> i["lut_log"]) + 1) * 4 for i in infos if i["learned"])
    q_s = res["learned"]["ms"] / 1e3
    achieved = sec["learned"] * 64.0 * nq / q_s / 1e9
    out = {"learned_Mq_s": res["learned"]["Mq_s"], "plain_Mq_s": res["plain"]["Mq_s"],
           "speedup": round(res["learned"]["Mq_s"] / res["plain"]["Mq_s"], 3),
           "identical_answers": same, "answers_digest_match": digest_ok,
           "sectors_per_query": sec,
           "roofline": {"bound": "hbm", "kernel": "k_lsel_query<true>", "achieved": round(achieved, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "bytes_per_query": round(sec["learned"] * 64.0, 1)},
           "build_and_train_s": round(build_s, 4), "build_and_train_cold_s": round(build_cold_s, 4),
           "learned_vectors": nl,
           "trained_vectors": sum(1 for i in infos if i["learned"] and i["trained"]),
           "placed_only": placed, "model_bytes": model_bytes,
           "training": {"loss_placed_mean": round(float(np.mean([i["loss_init"] for i in infos if i["learned"]] or [0])), 1),
                        "loss_trained_mean": round(float(np.mean([i["loss_final"] for i in infos if i["learned"]] or [0])), 1),
                        "vectors_loss_lowered": sum(1 for i in infos if i["learned"] and i["loss_final"] < i["loss_init"]),
                        "note": "Adam lowers the mean squared error of every model; the query reads a fixed "
                                "three-block window once eps <= 512, so trained and placed-only models read "
                                "the same sectors"},
           "eps_max_learned": max([i["eps"] for i in infos if i["learned"]] or [0]),
           "config": {"workload": "learned select, %d bitvectors x 2^%d bits, %d uniform queries"
                                  % (nvec, args.lsel_log2bits, nq), "tol": args.lsel_tol,
                      "max_gap": args.lsel_gap, "epochs": args.lsel_epochs}}
    if not full:
        out["per_vector"] = [{"d": round(float(dens[b]), 4), "learned": infos[b]["learned"],
                              "trained": infos[b]["trained"], "nknots": infos[b]["nknots"],
                              "eps": infos[b]["eps"], "eps_init": infos[b]["eps_init"],
                              "loss_init": round(infos[b]["loss_init"], 1),
                              "loss_final": round(infos[b]["loss_final"], 1)} for b in range(nvec)]
    ls.free()
    for a in dv + [dvq, djq, out_p, out_l]:
        a.free()
    ctx.close()
    return out, same and digest_ok is not False and (placed is None or placed["identical_answers"])


# --------------------------------------------------------------- C4
def bench_batch(pkg, args, full):
    """Batched level-wise trees: 4096 byte sequences, n_i = floor(2^U(10,22)),
    alphabets {4, 20, 256}; one hipGraph replay per step, against per-sequence
    plain launches; every sequence's words hashed against the reference levelWT."""
    import ctypes as C
    W = pkg.workloads
    nseq = args.batch_seqs
    log("generating %d sequences" % nseq)
    n, sigma, sym, off = W.c4_batch(nseq, float(args.batch_log2max))
    total = int(n.sum())
    ctx = pkg.Context(0)
    d = ctx.to_device(sym)
    t0 = time.perf_counter()
    if args.batch_graph_only:  # one group per alphabet width, no tuning builds (for profiles)
        plan, tune_ms = np.array([0xFF] + [24 - w for w in range(1, 9)], np.uint8), None
    else:
        plan, tune_ms = pkg.Batch.tune(ctx, d, n, sigma, off, reps=2)
    tune_s = time.perf_counter() - t0
    B = pkg.Batch(ctx, d, n, sigma, off, plan=plan)
    L = np.array([pkg.levels_of_sigma(int(x)) for x in sigma])
    log("batch: %d symbols, plan %s" % (total, list(plan)))
    B.run()
    ctx.set_blocking(False)
    for _ in range(args.warmup):
        B.run()
    ctx.sync()
    ms = timed(ctx, B.run, args.steps)
    ctx.sync()
    lib = pkg.lib()
    wo = np.concatenate([[0], np.cumsum([lib.wtg_levels_words(int(k), int(l)) + 1
                                         for k, l in zip(n, L)])]).astype(np.uint64)
    dw = ctx.empty(int(wo[-1]), np.uint64)

    if args.batch_graph_only:
        return {"value": round(total / (ms / 1e3) / 1e9, 4), "unit": "Gsym/s", "ms_per_step": round(ms, 4),
                "symbols": total, "note": "graph replays only: no tuner, per-sequence builds or digests"}, True

    def per_seq():
        for i in range(nseq):
            if n[i]:
                pkg._check(lib.wtg_levels_build(ctx.h, C.c_void_p(d.ptr.value + int(off[i])), 1, int(n[i]),
                                                int(L[i]), C.c_void_p(dw.ptr.value + 8 * int(wo[i])), None, 0))
    per_seq()
    ctx.sync()
    ms_seq = timed(ctx, per_seq, 1)
    ctx.sync()
    ctx.set_blocking(True)
    # parity: every sequence of the batch, in the WT() layout, hashed in order
    # against the reference levelWT's words of the same sequences
    wd = golden("workload_digests.json").get("c4")
    digest_ok = None
    if wd and nseq == W.C4_SEQS and args.batch_log2max == 22:
        h = hashlib.sha256()
        for i in range(nseq):
            h.update(B.extract(i).tobytes())
        digest_ok = h.hexdigest() == wd["words_sha256"]
    allw = dw.get()
    ok = True
    rng = np.random.default_rng(1)
    for i in list(rng.choice(nseq, 24, replace=False)) + [int(np.argmax(n))]:
        e = allw[int(wo[i]):int(wo[i]) + lib.wtg_levels_words(int(n[i]), int(L[i]))]
        ok &= bool(np.array_equal(B.extract(int(i)), e))
    del allw
    B_a = float((n.astype(np.float64) * (2 * L + L / 8.0)).sum())
    out = {"value": round(total / (ms / 1e3) / 1e9, 4), "unit": "Gsym/s", "ms_per_step": round(ms, 4),
           "symbols": total, "build_roofline": {"B_a_bytes": B_a, "frac": frac(B_a, ms)},
           "words_digest_match": digest_ok, "parity_vs_per_sequence": ok,
           "per_sequence_launches": {"ms": round(ms_seq, 3), "Gsym_s": round(total / (ms_seq / 1e3) / 1e9, 4)},
           "speedup_vs_per_sequence": round(ms_seq / ms, 2),
           "autotune": {"cap_log_by_width": {int(w): int(plan[w]) for w in sorted(set(L.tolist()))},
                        "tune_s": round(tune_s, 3)},
           "config": {"workload": "batched levelWT, %d sequences, n_i=floor(2^U(10,%d)), alphabets {4,20,256}"
                                  % (nseq, args.batch_log2max)}}
    if not full and tune_ms is not None:
        out["autotune"]["candidates_ms"] = {int(w): [round(float(x), 3) for x in tune_ms[w]]
                                            for w in sorted(set(L.tolist()))}
    if not args.no_cpu:
        from oracle import oracle
        ref = oracle.RefWT("levelWT", "u8")
        acc_n, acc_t, i = 0, 0.0, 0
        for i in range(nseq):
            if acc_n >= (1 << 26):
                break
            s_i = sym[int(off[i]):int(off[i]) + int(n[i])]
            if ref.ok:
                acc_t += float(ref.time(s_i, int(sigma[i]), rounds=1).min())
            else:
                t0 = time.perf_counter()
                oracle.levels_serial(s_i, int(L[i]))
                acc_t += time.perf_counter() - t0
            acc_n += int(n[i])
        out["cpu_baseline"] = {"value": round(acc_n / acc_t / 1e9, 5), "unit": "Gsym/s",
                               "cores": int(ref.workers()) if ref.ok else 1,
                               "kind": "reference" if ref.ok else "port",
                               "sample": "the first %d sequences' symbols (%d), one levelWT WT() each" % (i, acc_n)}
    B.free(); d.free(); dw.free(); ctx.close()
    return out, ok and digest_ok is not False


# --------------------------------------------------------------- C3
def bench_fm(pkg, args, full):
    """FM-index over synthetic DNA: GPU suffix array + BWT (checked against the
    reference DC3 suffix array's digest), wt_huff build on the GPU (timed per
    build, input resident), then batched backward-search counts of length-20
    patterns (half exact substrings, half with 1-2 substitutions)."""
    W = pkg.workloads
    n = 1 << args.fm_log2n
    log("generating %d DNA bytes" % n)
    t = W.c3_text(n)
    pats = W.c3_patterns(t, args.fm_patterns)
    wd = golden("workload_digests.json").get("c3") if n == W.C3_N and args.fm_patterns == W.C3_PATTERNS else None
    ctx = pkg.Context(0)
    dt = ctx.to_device(t)
    pkg.suffix_array_gpu(ctx, dt, n).free()
    ctx.sync()
    t0 = time.perf_counter()
    dsa = pkg.suffix_array_gpu(ctx, dt, n)
    ctx.sync()
    gpu_sa_s = time.perf_counter() - t0
    dbwt = pkg.bwt_from_sa_gpu(ctx, dt, dsa, n)
    par = {}
    host_sa = None
    if wd:
        par["sa_digest_match"] = sha(dsa.get()) == wd["sa_sha256"]
        par["bwt_digest_match"] = sha(dbwt.get()) == wd["bwt_sha256"]
    else:  # smaller texts: the host SA-IS
        host_sa = pkg.suffix_array(t)
        par["sa_matches_host_sais"] = bool(np.array_equal(dsa.get(), host_sa))
    dsa.free(); dt.free()
    for _ in range(max(args.warmup, 1)):
        ctx.huff_build(dbwt).free()
    build_ms = timed(ctx, lambda: ctx.huff_build(dbwt).free(), args.steps)
    ctx.kernel_timing_reset()
    ctx.kernel_timing(True)
    for _ in range(args.steps):
        ctx.huff_build(dbwt).free()
    kt = {k: ctx.kernel_time(k) for k in ("huff_hist", "huff_count", "huff_scan", "huff_emit", "rank_select")}
    ctx.kernel_timing(False)
    wt = ctx.huff_build(dbwt)
    dp = ctx.to_device(pats.reshape(-1))
    do = ctx.empty(pats.shape[0], np.uint64)
    lib = pkg.lib()
    cnt = lambda: pkg._check(lib.wtg_fm_count_batch(ctx.h, wt.h, dp.ptr, W.C3_M, pats.shape[0], do.ptr))
    cnt()
    q_ms = timed(ctx, cnt, args.steps)
    got = do.get()
    if wd:
        par["counts_digest_match"] = sha(got) == wd["counts_sha256"]
    else:  # every pattern drawn from the text occurs at least once
        par["exact_patterns_found"] = bool((got[0::2] > 0).all())
    bits = int(wt.shape.bv_size)
    B_c = n * 1.0 + 2.0 * bits + bits / 8.0   # n*w + sum_l a_l*2w + n*H/8 (w = 1, sum_l a_l = bits)
    emit_ms = kt["huff_emit"][0] / max(kt["huff_emit"][1], 1)
    out = {"value": round(n / (build_ms / 1e3) / 1e9, 4), "unit": "Gsym/s", "build_ms": round(build_ms, 4),
           "bits": bits, "mean_code_len": round(bits / n, 4),
           "build_roofline": {"B_c_bytes": B_c, "frac": frac(B_c, build_ms)},
           "kernels_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in kt.items()},
           "emit_GBps": round((n + bits / 8.0) / (emit_ms / 1e3) / 1e9, 1) if emit_ms else None,
           "count_ms": round(q_ms, 4), "count_Mpat_s": round(pats.shape[0] / (q_ms / 1e3) / 1e6, 2),
           "gpu_sa_s": round(gpu_sa_s, 4), "parity": par,
           "config": {"workload": "FM-index over 2^%d-byte order-5 Markov DNA, %d length-20 patterns"
                                  % (args.fm_log2n, pats.shape[0])}}
    if not args.no_cpu:
        from oracle import oracle
        bwt_h = dbwt.get()
        m = min(n, 1 << 24)
        t0 = time.perf_counter()
        oracle.Huff(bwt_h[:m])
        dtc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(m / dtc / 1e9, 5), "unit": "Gsym/s", "cores": 1, "kind": "port",
                               "sample": "serial Huffman-shape + tree bits + rank9 + select restatement on the "
                                         "first %d BWT bytes (sdsl, the reference's wt_huff base, is absent)" % m}
        rs = oracle.RefSA()
        if rs.ok:
            ms_ = min(n, 1 << 25)
            t0 = time.perf_counter()
            rs.suffix_array(t[:ms_])
            dts = time.perf_counter() - t0
            out["cpu_baseline_suffix_array"] = {
                "value": round(ms_ / dts / 1e6, 3), "unit": "Msym/s", "cores": int(rs.workers()),
                "kind": "reference", "gpu_Msym_s": round(n / gpu_sa_s / 1e6, 3),
                "sample": "reference PBBS DC3 suffix array (suffixArray/parallelKS/pks.C, OpenMP) on the "
                          "first %d bytes of the text" % ms_}
    wt.free(); dbwt.free(); dp.free(); do.free(); ctx.close()
    return out, all(par.values())


# --------------------------------------------------------------- C1
def bench_bytes(pkg, args, full):
    """Byte wavelet tree (levelWT): n = 2^24 uint8 Zipf(1.1) symbols over a seeded
    byte permutation, 8 levels + rank directory per step; 10^6 uniform rank(c, i)
    queries; words and answers hashed against the reference / brute force."""
    W = pkg.workloads
    n, L = W.C1_N, W.C1_LEVELS
    s = W.c1_input()
    ctx = pkg.Context(0)
    d = ctx.to_device(s)
    lib = pkg.lib()
    dw = ctx.empty(lib.wtg_levels_words(n, L), np.uint64)
    D = ctx.empty(lib.wtg_rank9_words(n * L), np.uint64)

    def step():
        ctx.levels_build(d, n, L, dw)
        pkg._check(lib.wtg_rank9_build(ctx.h, dw.ptr, n * L, D.ptr))
    ctx.set_blocking(False)
    for _ in range(args.warmup):
        step()
    ctx.sync()
    ms = timed(ctx, step, args.steps)
    ctx.sync()
    ctx.set_blocking(True)
    fix = golden("full_size_digests.json").get("c1_zipf_bytes")
    par = {"words_digest_match": sha(dw.get()) == fix["words_sha256"]} if fix else {}
    wt = ctx.wtint_build(d, n, L)
    qi, qc = W.c1_queries()
    di, dc, do = ctx.to_device(qi), ctx.to_device(qc), ctx.empty(qi.size, np.uint64)
    rq = lambda: pkg._check(lib.wtg_wtint_rank_batch(ctx.h, wt.h, di.ptr, dc.ptr, qi.size, do.ptr))
    rq()
    q_ms = timed(ctx, rq, args.steps)
    wd = golden("workload_digests.json").get("c1")
    if wd:
        par["rank_answers_digest_match"] = sha(do.get()) == wd["rank_answers_sha256"]
    B_a = n * (2 * L * 1 + L / 8.0)
    out = {"value": round(n / (ms / 1e3) / 1e9, 4), "unit": "Gsym/s", "ms_per_step": round(ms, 4),
           "build_roofline": {"B_a_bytes": B_a, "frac": frac(B_a, ms)},
           "rank_queries_Mq_s": round(qi.size / (q_ms / 1e3) / 1e6, 1), "parity": par,
           "config": {"workload": "levelWT bytes, n=2^24 Zipf(1.1) over a byte permutation, 8 levels + rank9"}}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_levels(s, L, n, 3, "u8")
        out["cpu_baseline"]["note"] = "tree only (no rank directory)"
    for a in (d, dw, D, di, dc, do):
        a.free()
    wt.free()
    ctx.close()
    return out, all(par.values())


def single_line(name, res, args):
    """The one-workload line of --workload (metric/value/unit first)."""
    metrics = {"bytes": ("byte WT build Gsym/s", "value", "Gsym/s", "u8"),
               "batch": ("batched WT build Gsym/s", "value", "Gsym/s", "u8"),
               "fm": ("wt_huff build over a BWT Gsym/s; FM count Mpat/s", "value", "Gsym/s", "u8"),
               "lsel": ("select Mq/s learned vs plain (bit-exact)", "learned_Mq_s", "Mq/s", "u64")}
    m, key, unit, dt = metrics[name]
    out = {"metric": m, "value": res[key], "unit": unit, "n_gpus": 1, "steps": args.steps,
           "higher_is_better": True, "dtype": dt, "data": "synthetic"}
    out.update({k: v for k, v in res.items() if k not in ("value",)})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--levels", type=int, default=20)
    ap.add_argument("--cpu-log2n", type=int, default=26)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="the wt_int line without the learned-select, C1, C3 and C4 blocks")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device-to-device twice (a plain copy of known size, "
                         "the reference point for HBM byte counters)")
    ap.add_argument("--no-geometric", dest="geometric", action="store_false",
                    help="skip the geometric(1e-4) second input")
    ap.add_argument("--queries", type=int, default=10_000_000,
                    help="rank and select queries on the built tree (0: skip)")
    ap.add_argument("--workload", default=None, choices=["wt_int", "bytes", "lsel", "batch", "fm"],
                    help="print one workload's line only")
    ap.add_argument("--lsel-log2bits", type=int, default=26)
    ap.add_argument("--batch-seqs", type=int, default=4096)
    ap.add_argument("--fm-log2n", type=int, default=27)
    ap.add_argument("--fm-patterns", type=int, default=1_000_000)
    ap.add_argument("--batch-log2max", type=int, default=22)
    ap.add_argument("--batch-graph-only", action="store_true",
                    help="C4: time the graph replays only (one group per alphabet width, no tuner, "
                         "no per-sequence builds or digests), for kernel profiles")
    ap.add_argument("--lsel-queries", type=int, default=100_000_000)
    ap.add_argument("--lsel-tol", type=int, default=480, help="learned select: knot placement tolerance (bits)")
    ap.add_argument("--lsel-gap", type=int, default=4096, help="learned select: longest knot interval (ones)")
    ap.add_argument("--lsel-epochs", type=int, default=50, help="learned select: Adam steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every workload of this path runs on a single "
              "MI355X; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        return 2
    pkg = load_pkg()
    subs = {"lsel": bench_lsel, "bytes": bench_bytes, "batch": bench_batch, "fm": bench_fm}
    if args.workload in subs:
        res, ok = subs[args.workload](pkg, args, False)
        print(json.dumps(single_line(args.workload, res, args)), flush=True)
        return 0 if ok else 1
    out, ok = bench_wt_int(pkg, args, True)
    if args.workload is None and not args.headline_only:
        # the other configurations, each at its full size with the default step counts
        sub_args = argparse.Namespace(**vars(args))
        sub_args.steps, sub_args.warmup = 10, 3
        for key, name in (("learned_select", "lsel"), ("c1_bytes", "bytes"), ("c3_fm", "fm"),
                          ("c4_batch", "batch")):
            log(key)
            t0 = time.perf_counter()
            res, sok = subs[name](pkg, sub_args, True)
            res["wall_s"] = round(time.perf_counter() - t0, 1)
            out[key] = res
            ok = ok and sok
    print(json.dumps(out), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
