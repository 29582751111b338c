"""The C++ drop-in layer, driven the way reference code calls it: WT() from
wavelet/levelWT/WT.h (byte, INT and INT+LONG builds), sdsl-shaped
rank_support_v / select_support_mcl<1> / <0> and wt_int<>.  The driver
(tests/cpp/dropin_driver.cpp, built by build()) writes its answers; they must
equal the oracle's and brute force bit for bit."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")


def run_driver(variant, outdir):
    exe = os.path.join(BUILD, "dropin_" + variant)
    if not os.path.exists(exe):
        pytest.fail("drop-in driver not built: run __graft_entry__.build()")
    print("driver", variant, flush=True)
    r = subprocess.run([exe, outdir], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "dropin ok" in r.stdout


def load(path, dt=np.uint64):
    return np.fromfile(path, dtype=dt)


@pytest.mark.parametrize("variant,dt", [("u8", np.uint8), ("u32", np.uint32), ("u64", np.uint64)])
def test_dropin_driver(wsfm, orc, variant, dt):
    with tempfile.TemporaryDirectory() as d:
        run_driver(variant, d)
        # WT()
        n = 100003
        sigma = 256 if variant == "u8" else 1000
        s = wsfm.gen_uniform(11, n, sigma, dt)
        L = orc.levels_of_sigma(sigma)
        ow, ond = orc.levels_serial(s, L)
        assert np.array_equal(load(d + "/wt_words.bin"), ow)
        nodes = load(d + "/wt_nodes.bin").reshape(2 * sigma, 5)
        reached = np.nonzero(ond[:, 5])[0]
        assert np.array_equal(nodes[reached], ond[reached, :5])
        # rank / select
        nbits = 1000003
        nw = (nbits + 63) // 64
        w = wsfm.gen_uniform(12, nw, (1 << 64) - 1, np.uint64)
        w[-1] &= np.uint64((1 << (nbits & 63)) - 1)
        bits = np.unpackbits(w.view(np.uint8), bitorder="little")[:nbits].astype(np.int64)
        pref = np.concatenate([[0], np.cumsum(bits)])
        idx = np.concatenate([np.arange(0, nbits + 1, 7), [nbits]])
        assert np.array_equal(load(d + "/rank.bin"), pref[idx].astype(np.uint64))
        p1 = np.nonzero(bits == 1)[0]
        p0 = np.nonzero(bits == 0)[0]
        assert np.array_equal(load(d + "/sel1.bin"), p1[::5].astype(np.uint64))
        assert np.array_equal(load(d + "/sel0.bin"), p0[::5].astype(np.uint64))
        # two-bit patterns `10` / `01`: rank over whole words, select below size
        from test_oracle_rank_select import pattern_occ
        nbits = 200003
        nw = (nbits + 63) // 64
        w = wsfm.gen_uniform(14, nw, (1 << 64) - 1, np.uint64)
        w[-1] &= np.uint64((1 << (nbits & 63)) - 1)
        w[-1] |= np.uint64(1 << ((nbits - 1) & 63))
        idx = np.concatenate([np.arange(0, nbits + 1, 3), [nbits]])
        rp = load(d + "/rank_pat.bin")
        got = {10: np.concatenate([rp[0:-2:2], rp[-2:-1]]), 1: np.concatenate([rp[1:-2:2], rp[-1:]])}
        for pat, name in ((10, "sel10"), (1, "sel01")):
            occ = pattern_occ(w, nw * 64, pat)
            pref = np.concatenate([[0], np.cumsum(occ)])
            assert np.array_equal(got[pat], pref[idx].astype(np.uint64)), pat
            pos = np.nonzero(occ[:nbits])[0]
            sel = load(d + "/" + name + ".bin")
            assert sel[0] == pos.size and np.array_equal(sel[1:], pos[::2].astype(np.uint64)), pat
        occ10 = pattern_occ(w, nw * 64, 10)
        assert occ10[nbits] == 1  # counted by the whole-word directory, not by select
        pw = orc.pattern_words(w, nbits, 10, clip=False)
        assert np.array_equal(load(d + "/rank10_dir.bin"), orc.rank9_build(pw, nbits))
        # wt_int
        m = 50000
        sv = wsfm.gen_uniform(13, m, 5000, np.uint32)
        res = load(d + "/wtint.bin")
        Lw = int(max(1, int(sv.max()))).bit_length()
        assert res[0] == Lw and res[1] == len(np.unique(sv))
        rows = res[2:].reshape(-1, 5)
        for k, i in enumerate(range(0, m, 97)):
            c = int(sv[(i * 31) % m])
            acc, rk, isr, isc, sel = (int(x) for x in rows[k])
            assert acc == sv[i]
            assert rk == int((sv[:i] == c).sum())
            assert isc == sv[i] and isr == int((sv[:i] == sv[i]).sum())
            occ = np.nonzero(sv == c)[0]
            j = 1 + (i % 7)
            assert sel == (occ[j - 1] if j <= occ.size else orc.NONE)
        # lex_count / lex_smaller_count (wt_int_par.hpp:712-780) by brute force
        lx = load(d + "/wtint_lex.bin").reshape(-1, 8)
        top = 1 << Lw
        for i, j, c, rk, sm, gr, rk2, sm2 in (tuple(int(x) for x in r) for r in lx):
            seg = sv[i:j].astype(np.int64)
            if c >= top:
                assert (rk, sm, gr) == (0, j - i, 0)
                assert (rk2, sm2) == (0, j)
            else:
                assert (rk, sm, gr) == (int((sv[:i] == c).sum()), int((seg < c).sum()), int((seg > c).sum()))
                assert (rk2, sm2) == (int((sv[:j] == c).sum()), int((sv[:j] < c).sum()))
        # range_search_2d: count, and the reported points in (value, index) order
        rs = load(d + "/wtint_range.bin")
        p = 0
        while p < rs.size:
            lb, rb, vlb, vrb, cnt, npt = (int(x) for x in rs[p:p + 6])
            p += 6
            pts = rs[p:p + 2 * npt].reshape(-1, 2)
            p += 2 * npt
            idx = np.arange(lb, rb + 1)
            val = sv[lb:rb + 1].astype(np.int64)
            sel = (val >= vlb) & (val <= vrb)
            assert cnt == int(sel.sum())
            if npt:
                exp = sorted(zip(val[sel].tolist(), idx[sel].tolist()))
                assert [(int(b), int(a)) for a, b in pts] == exp
        # interval_symbols (wt_int_par.hpp:684-702): distinct values ascending with ranks
        isy = load(d + "/wtint_isym.bin")
        p = 0
        while p < isy.size:
            i, j, k = (int(x) for x in isy[p:p + 3])
            p += 3
            got = isy[p:p + 3 * k].reshape(-1, 3)
            p += 3 * k
            vals = np.unique(sv[i:j])
            assert k == vals.size
            for (c, ri, rj), v in zip(got.tolist(), vals.tolist()):
                assert c == v and ri == int((sv[:i] == v).sum()) and rj == int((sv[:j] == v).sum())
        # node interface (wt_int_par.hpp:915-1085): leaves by expand from the root,
        # and ranges followed down the path of c
        lv = load(d + "/wtint_leaves.bin").reshape(-1, 2)
        vals, cnts = np.unique(sv, return_counts=True)
        assert np.array_equal(lv[:, 0], vals.astype(np.uint64)) and np.array_equal(lv[:, 1], cnts.astype(np.uint64))
        for i, j, c, n1, n2, n3, vsz, vsym in load(d + "/wtint_expand.bin").reshape(-1, 8).tolist():
            assert vsym == c and vsz == int((sv == c).sum())
            assert n1 == n2 == int((sv[i:j + 1] == c).sum())
            assert n3 == vsz
        it = load(d + "/wtint_iter.bin")
        assert it[-1] == m and np.array_equal(it[:-1], sv.astype(np.uint64))
        # FMIndex() + wt_huff<>
        n = 200003
        t = wsfm.gen_dna(14, n)
        sa = wsfm.suffix_array(t)
        bwt = orc.bwt_from_sa(t, sa)
        assert np.array_equal(load(d + "/fm_bwt.bin", np.uint8), bwt)
        res = load(d + "/huff.bin")
        assert res[0] == len(np.unique(bwt))
        rows = res[1:].reshape(-1, 5)
        # interval_symbols (wt_pc_par.hpp:579-617): tree order (left child first,
        # from the oracle's shape) from three positions on, text order for two
        H = orc.Huff(bwt).H
        order = []
        stack = [0]
        while stack:
            v = stack.pop()
            if H.child0[v] < 0:
                order.append(int(H.sym[v]))
            else:
                stack += [H.child1[v], H.child0[v]]
        hsy = load(d + "/huff_isym.bin")
        p = 0
        while p < hsy.size:
            i, j, k = (int(x) for x in hsy[p:p + 3])
            p += 3
            got = hsy[p:p + 3 * k].reshape(-1, 3).tolist()
            p += 3 * k
            seg = bwt[i:j]
            if j - i == 2 and seg[0] != seg[1]:
                exp_syms = [int(seg[0]), int(seg[1])]
            else:
                present = set(np.unique(seg).tolist())
                exp_syms = [c for c in order if c in present]
            assert [g[0] for g in got] == exp_syms, (i, j)
            for c, ri, rj in got:
                assert ri == int((bwt[:i] == c).sum()) and rj == int((bwt[:j] == c).sum())
        cs = [65, 67, 71, 84, 78, 0]
        for k, i in enumerate(range(0, n, 101)):
            c = cs[k % 6]
            acc, rk, isr, isc, sel = (int(x) for x in rows[k])
            assert acc == bwt[i] and isc == bwt[i]
            if k % 40 == 0:
                assert rk == int((bwt[:i] == c).sum())
                assert isr == int((bwt[:i] == bwt[i]).sum())
                occ = np.nonzero(bwt == c)[0]
                j = 1 + (i % 5000)
                assert sel == (occ[j - 1] if j <= occ.size else orc.NONE)
        # iterators, node interface, symbol_gte / symbol_lte (wt_pc_par.hpp:740-909)
        assert np.array_equal(load(d + "/huff_iter.bin", np.uint8), bwt)
        hl = load(d + "/huff_leaves.bin").reshape(-1, 2)
        vals, cnts = np.unique(bwt, return_counts=True)
        assert sorted(map(tuple, hl.tolist())) == list(zip(vals.tolist(), cnts.tolist()))
        assert [int(x) for x in hl[:, 0]] == order   # leaves in tree order, left child first
        hp = load(d + "/huff_nodes.bin")
        present = set(vals.tolist())
        p = 0
        for c in range(256):
            g1, g, l1, lo = (int(x) for x in hp[p:p + 4])
            p += 4
            ge = [x for x in range(c, 256) if x in present]
            le = [x for x in range(c, -1, -1) if x in present]
            assert (g1, g) == ((1, ge[0]) if ge else (0, 0))
            assert (l1, lo) == ((1, le[0]) if le else (0, 0))
            if c in present:
                cc, plen, code, i, j, rsz, fsz, leaf, sym = (int(x) for x in hp[p:p + 9])
                p += 9
                assert cc == c and leaf == 1 and sym == c
                assert plen == H.path_len[c]
                assert code == int("".join(str((H.path[c] >> k) & 1) for k in range(plen)) or "0", 2)
                assert rsz == int((bwt[i:j + 1] == c).sum()) and fsz == int((bwt == c).sum())
        assert p == hp.size
        pats = wsfm.gen_patterns(15, t, 12, 300)
        assert np.array_equal(load(d + "/fm_count.bin"), orc.sa_count(t, sa, pats))
