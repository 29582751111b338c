"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.  It loads ``_build/liboracle.so`` (the C
restatement in wtg_oracle.c) and, when present, the compiled reference builders
``_ref/libref_<alg>_<sym>.so`` (reference wavelet/<alg>/WT.C + ref_harness.cpp).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
U64P = C.POINTER(C.c_uint64)
NONE = (1 << 64) - 1


def build(with_ref=None):
    """Compile the oracle (and the reference builders when /root/reference exists)."""
    targets = ["all"]
    ref_root = "/root/reference"
    if with_ref is None:
        with_ref = os.path.isdir(os.path.join(ref_root, "wavelet"))
    if with_ref:
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE, "-j8", "REF=" + ref_root] + targets, check=True)


class OrcSelect(C.Structure):
    _fields_ = [("nbits", C.c_uint64), ("arg_cnt", C.c_uint64), ("sb", C.c_uint64),
                ("n_long", C.c_uint64), ("logn", C.c_uint32), ("logn4", C.c_uint32),
                ("superblock", U64P), ("mini", C.POINTER(C.c_uint32)),
                ("mini_width", C.POINTER(C.c_uint8)), ("long_index", C.POINTER(C.c_int64)),
                ("long_width", C.POINTER(C.c_uint8)), ("long_pool", U64P)]


class OrcHuff(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("sigma", C.c_uint32), ("bv_size", C.c_uint64),
                ("bv_pos", C.c_uint64 * 511), ("freq", C.c_uint64 * 511),
                ("child0", C.c_int32 * 511), ("child1", C.c_int32 * 511),
                ("parent", C.c_int32 * 511), ("sym", C.c_int32 * 511),
                ("c_to_leaf", C.c_int32 * 256), ("path", C.c_uint64 * 256),
                ("path_len", C.c_uint32 * 256)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            build(with_ref=False)
        L = C.CDLL(path)
        vp, u64, i32, u32 = C.c_void_p, C.c_uint64, C.c_int, C.c_uint32
        L.orc_remap.restype = u64
        L.orc_remap.argtypes = [vp, i32, u64]
        L.orc_levels_serial.argtypes = [vp, i32, u64, u32, vp, vp]
        L.orc_levels_closed.argtypes = [vp, i32, u64, u32, vp]
        L.orc_wt_query.restype = u64
        L.orc_wt_query.argtypes = [vp, vp, u64, u64]
        L.orc_rank9_entries.restype = u64
        L.orc_rank9_entries.argtypes = [u64]
        L.orc_rank9_build.argtypes = [vp, u64, vp]
        L.orc_pattern_words.argtypes = [vp, u64, u32, i32, vp]
        L.orc_rank9_rank.restype = u64
        L.orc_rank9_rank.argtypes = [vp, vp, u64]
        L.orc_select_plan.argtypes = [vp, u64, i32, C.POINTER(OrcSelect)]
        L.orc_select_fill.argtypes = [vp, u64, i32, C.POINTER(OrcSelect)]
        L.orc_select_query.restype = u64
        L.orc_select_query.argtypes = [C.POINTER(OrcSelect), vp, i32, u64]
        L.orc_rank9_rank_batch.restype = None
        L.orc_rank9_rank_batch.argtypes = [vp, vp, u64, vp, u64, vp]
        L.orc_select_query_batch.restype = None
        L.orc_select_query_batch.argtypes = [C.POINTER(OrcSelect), vp, i32, vp, u64, vp]
        L.orc_select_brute.restype = u64
        L.orc_select_brute.argtypes = [vp, u64, i32, u64]
        L.orc_wtint_rank.restype = u64
        L.orc_wtint_rank.argtypes = [vp, vp, u64, u32, u64, u64]
        L.orc_wtint_lex_count.restype = None
        L.orc_wtint_lex_count.argtypes = [vp, vp, u64, u32, u64, u64, u64, vp]
        L.orc_wtint_lex_smaller.restype = None
        L.orc_wtint_lex_smaller.argtypes = [vp, vp, u64, u32, u64, u64, vp]
        L.orc_wtint_range_count.restype = u64
        L.orc_wtint_range_count.argtypes = [vp, vp, u64, u32, u64, u64, u64, u64]
        L.orc_wtint_access.restype = u64
        L.orc_wtint_access.argtypes = [vp, vp, u64, u32, u64]
        L.orc_wtint_select.restype = u64
        L.orc_wtint_select.argtypes = [vp, vp, C.POINTER(OrcSelect), C.POINTER(OrcSelect),
                                       u64, u32, u64, u64]
        L.orc_huff_shape.argtypes = [vp, C.POINTER(OrcHuff)]
        L.orc_huff_bits.argtypes = [vp, u64, C.POINTER(OrcHuff), vp]
        L.orc_huff_rank.restype = u64
        L.orc_huff_rank.argtypes = [C.POINTER(OrcHuff), vp, vp, u64, u32]
        L.orc_huff_access.restype = u64
        L.orc_huff_access.argtypes = [C.POINTER(OrcHuff), vp, vp, u64]
        L.orc_huff_select.restype = u64
        L.orc_huff_select.argtypes = [C.POINTER(OrcHuff), vp, C.POINTER(OrcSelect),
                                      C.POINTER(OrcSelect), vp, u64, u64, u32]
        L.orc_bwt_from_sa.argtypes = [vp, vp, u64, vp]
        L.orc_sa_brute.argtypes = [vp, u64, vp]
        L.orc_fm_count.argtypes = [vp, u64, vp, u32, u64, vp]
        L.orc_sa_count.argtypes = [vp, vp, u64, vp, u32, u64, vp]
        L.orc_log2up.restype = C.c_int
        L.orc_log2up.argtypes = [u64]
        L.orc_pbbs_hash.restype = C.c_uint32
        L.orc_pbbs_hash.argtypes = [C.c_uint32]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def log2up(x):
    return lib().orc_log2up(int(x))


def levels_of_sigma(sigma):
    """L = max(1, log2Up(sigma)) (wavelet/levelWT/WT.C:53)."""
    return max(1, log2up(sigma))


def nwords(nbits):
    return (int(nbits) + 63) // 64


def remap(s):
    s = np.ascontiguousarray(s).copy()
    sigma = lib().orc_remap(_p(s), s.dtype.itemsize, s.size)
    return s, int(sigma)


def levels_serial(s, L, with_nodes=True):
    s = np.ascontiguousarray(s)
    n = s.size
    words = np.zeros(max(1, nwords(n * L)), dtype=np.uint64)
    nodes = np.zeros((1 << L, 6), dtype=np.uint64) if with_nodes else None
    rc = lib().orc_levels_serial(_p(s), s.dtype.itemsize, n, L, _p(words),
                                 _p(nodes) if with_nodes else None)
    assert rc == 0
    return words[:nwords(n * L)], nodes


def levels_closed(s, L):
    s = np.ascontiguousarray(s)
    n = s.size
    words = np.zeros(max(1, nwords(n * L)), dtype=np.uint64)
    assert lib().orc_levels_closed(_p(s), s.dtype.itemsize, n, L, _p(words)) == 0
    return words[:nwords(n * L)]


def wt_query(nodes, words, index, sigma):
    return lib().orc_wt_query(_p(nodes), _p(words), int(index), int(sigma))


def padded(words):
    """words + one zero padding word (sdsl's allocation always has one)."""
    out = np.zeros(words.size + 1, dtype=np.uint64)
    out[:words.size] = words
    return out


def rank9_build(words, nbits):
    D = np.zeros(lib().orc_rank9_entries(nbits), dtype=np.uint64)
    w = padded(words)
    lib().orc_rank9_build(_p(w), int(nbits), _p(D))
    return D


def pattern_words(words, nbits, pattern, clip):
    """argument words of the two-bit pattern `10` (pattern=10) or `01` (pattern=1)"""
    w = np.ascontiguousarray(words, dtype=np.uint64)
    out = np.zeros(max(1, (int(nbits) + 63) // 64), np.uint64)
    lib().orc_pattern_words(_p(w), int(nbits), int(pattern), int(bool(clip)), _p(out))
    return out


def rank9(D, words, idx, nbits=None):
    """rank(idx) per rank_support_v<1,1>; with nbits, idx > nbits gives NONE."""
    w = padded(words)
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    out = np.empty(idx.size, np.uint64)
    nb = w.size * 64 if nbits is None else int(nbits)
    lib().orc_rank9_rank_batch(_p(D), _p(w), nb, _p(idx), idx.size, _p(out))
    return out


class Select:
    """Owns the arrays of one select_support_mcl content restatement."""

    def __init__(self, words, nbits, bit):
        self.words = padded(words)
        self.bit = int(bit)
        S = OrcSelect()
        L = lib()
        L.orc_select_plan(_p(self.words), int(nbits), self.bit, C.byref(S))
        sb, nl = S.sb, S.n_long
        self.superblock = np.zeros(max(sb, 1), np.uint64)
        self.mini = np.zeros(max(sb, 1) * 64, np.uint32)
        self.mini_width = np.zeros(max(sb, 1), np.uint8)
        self.long_index = np.full(max(sb, 1), -1, np.int64)
        self.long_width = np.zeros(max(sb, 1), np.uint8)
        self.long_pool = np.zeros(max(nl, 1) * 4096, np.uint64)
        S.superblock = self.superblock.ctypes.data_as(U64P)
        S.mini = self.mini.ctypes.data_as(C.POINTER(C.c_uint32))
        S.mini_width = self.mini_width.ctypes.data_as(C.POINTER(C.c_uint8))
        S.long_index = self.long_index.ctypes.data_as(C.POINTER(C.c_int64))
        S.long_width = self.long_width.ctypes.data_as(C.POINTER(C.c_uint8))
        S.long_pool = self.long_pool.ctypes.data_as(U64P)
        L.orc_select_fill(_p(self.words), int(nbits), self.bit, C.byref(S))
        self.S = S
        self.arg_cnt, self.sb, self.n_long = int(S.arg_cnt), int(sb), int(nl)
        self.logn, self.logn4 = int(S.logn), int(S.logn4)

    def query(self, i):
        return lib().orc_select_query(C.byref(self.S), _p(self.words), self.bit, int(i))

    def queries(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        out = np.empty(idx.size, np.uint64)
        lib().orc_select_query_batch(C.byref(self.S), _p(self.words), self.bit, _p(idx), idx.size, _p(out))
        return out


def select_brute(words, nbits, bit, i):
    w = padded(words)
    return lib().orc_select_brute(_p(w), int(nbits), int(bit), int(i))


class WtInt:
    """Oracle wt_int: levels (closed form == serialWT), rank9, select1/select0."""

    def __init__(self, s, L):
        self.s = np.ascontiguousarray(s)
        self.n, self.L = self.s.size, int(L)
        self.words, _ = levels_serial(self.s, self.L, with_nodes=False)
        self.nbits = self.n * self.L
        self.D = rank9_build(self.words, self.nbits)
        self.S1 = Select(self.words, self.nbits, 1)
        self.S0 = Select(self.words, self.nbits, 0)
        self.wp = padded(self.words)

    def rank(self, i, c):
        return lib().orc_wtint_rank(_p(self.D), _p(self.wp), self.n, self.L, int(i), int(c))

    def access(self, i):
        return lib().orc_wtint_access(_p(self.D), _p(self.wp), self.n, self.L, int(i))

    def select(self, i, c):
        return lib().orc_wtint_select(_p(self.D), _p(self.wp), C.byref(self.S1.S),
                                      C.byref(self.S0.S), self.n, self.L, int(i), int(c))

    def lex_count(self, i, j, c):
        out = np.zeros(3, np.uint64)
        lib().orc_wtint_lex_count(_p(self.D), _p(self.wp), self.n, self.L, int(i), int(j), int(c), _p(out))
        return tuple(int(x) for x in out)

    def lex_smaller_count(self, i, c):
        out = np.zeros(2, np.uint64)
        lib().orc_wtint_lex_smaller(_p(self.D), _p(self.wp), self.n, self.L, int(i), int(c), _p(out))
        return tuple(int(x) for x in out)

    def range_count(self, lb, rb, vlb, vrb):
        return lib().orc_wtint_range_count(_p(self.D), _p(self.wp), self.n, self.L, int(lb), int(rb), int(vlb),
                                           int(vrb))


class Huff:
    def __init__(self, text):
        self.t = np.ascontiguousarray(text, dtype=np.uint8)
        self.n = self.t.size
        Cnt = np.bincount(self.t, minlength=256).astype(np.uint64)
        self.C = Cnt
        self.H = OrcHuff()
        lib().orc_huff_shape(_p(Cnt), C.byref(self.H))
        self.nbits = int(self.H.bv_size)
        self.words = np.zeros(max(1, nwords(self.nbits)), np.uint64)
        lib().orc_huff_bits(_p(self.t), self.n, C.byref(self.H), _p(self.words))
        self.words = self.words[:nwords(self.nbits)]
        self.D = rank9_build(self.words, self.nbits)
        self.S1 = Select(self.words, self.nbits, 1)
        self.S0 = Select(self.words, self.nbits, 0)
        self.wp = padded(self.words)

    def rank(self, i, c):
        return lib().orc_huff_rank(C.byref(self.H), _p(self.D), _p(self.wp), int(i), int(c))

    def access(self, i):
        return lib().orc_huff_access(C.byref(self.H), _p(self.D), _p(self.wp), int(i))

    def select(self, i, c):
        return lib().orc_huff_select(C.byref(self.H), _p(self.D), C.byref(self.S1.S),
                                     C.byref(self.S0.S), _p(self.wp), self.n, int(i), int(c))


def sa_brute(t):
    t = np.ascontiguousarray(t, dtype=np.uint8)
    sa = np.zeros(t.size, np.int64)
    lib().orc_sa_brute(_p(t), t.size, _p(sa))
    return sa


def bwt_from_sa(t, sa):
    t = np.ascontiguousarray(t, dtype=np.uint8)
    sa = np.ascontiguousarray(sa, dtype=np.int64)
    out = np.zeros(t.size, np.uint8)
    lib().orc_bwt_from_sa(_p(t), _p(sa), t.size, _p(out))
    return out


def fm_count(bwt, pats):
    bwt = np.ascontiguousarray(bwt, dtype=np.uint8)
    pats = np.ascontiguousarray(pats, dtype=np.uint8)
    out = np.zeros(pats.shape[0], np.uint64)
    lib().orc_fm_count(_p(bwt), bwt.size, _p(pats), pats.shape[1], pats.shape[0], _p(out))
    return out


def sa_count(t, sa, pats):
    t = np.ascontiguousarray(t, dtype=np.uint8)
    sa = np.ascontiguousarray(sa, dtype=np.int64)
    pats = np.ascontiguousarray(pats, dtype=np.uint8)
    out = np.zeros(pats.shape[0], np.uint64)
    lib().orc_sa_count(_p(t), _p(sa), t.size, _p(pats), pats.shape[1], pats.shape[0], _p(out))
    return out


# ----------------------------------------------------------------- reference
class RefWT:
    """The compiled reference builder wavelet/<alg>/WT.C (oracle/_ref), or None."""

    def __init__(self, alg, sym):
        path = os.path.join(HERE, "_ref", "libref_%s_%s.so" % (alg, sym))
        self.ok = os.path.exists(path)
        if not self.ok:
            return
        self.L = C.CDLL(path)
        self.L.ref_wt.restype = C.c_int
        self.L.ref_wt.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p]
        self.L.ref_wt_time.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_void_p]
        self.L.ref_symbol_bytes.restype = C.c_int
        self.L.ref_workers.restype = C.c_int
        self.symbytes = self.L.ref_symbol_bytes()

    def build(self, s, sigma):
        s = np.ascontiguousarray(s)
        assert s.dtype.itemsize == self.symbytes
        L = levels_of_sigma(sigma)
        words = np.zeros(max(1, nwords(s.size * L)), np.uint64)
        nodes = np.zeros((2 * sigma, 5), np.uint64)
        self.L.ref_wt(_p(s), s.size, sigma, _p(words), _p(nodes))
        return words[:nwords(s.size * L)], nodes

    def time(self, s, sigma, rounds=3):
        s = np.ascontiguousarray(s)
        t = np.zeros(rounds, np.float64)
        self.L.ref_wt_time(_p(s), s.size, sigma, rounds, _p(t))
        return t

    def workers(self):
        return self.L.ref_workers()


class RefSA:
    """The reference's PBBS DC3 suffix array (suffixArray/parallelKS/pks.C,
    compiled into oracle/_ref/libref_pks.so), or ok = False."""

    def __init__(self):
        path = os.path.join(HERE, "_ref", "libref_pks.so")
        self.ok = os.path.exists(path)
        if not self.ok:
            return
        self.L = C.CDLL(path)
        self.L.ref_suffix_array.restype = C.c_int
        self.L.ref_suffix_array.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        self.L.ref_sa_workers.restype = C.c_int

    def suffix_array(self, t):
        t = np.ascontiguousarray(t, dtype=np.uint8)
        sa = np.empty(t.size, np.int64)
        rc = self.L.ref_suffix_array(_p(t), t.size, _p(sa))
        assert rc == 0, "ref_suffix_array failed (%d)" % rc
        return sa

    def workers(self):
        return self.L.ref_sa_workers()


# ----------------------------------------------------------- learned select
def spline_interval(J, j):
    """Knot interval k of each argument j: J[k] <= j < J[k+1], clamped to [0, K-2]."""
    J = np.asarray(J, np.float64)
    return np.clip(np.searchsorted(J, np.asarray(j, np.float64), side="right") - 1, 0, J.size - 2)


def spline_predict(J, y, alpha, beta, j):
    """Continuous piecewise-linear select model (float64): alpha*j + beta plus the
    linear interpolation of residual knots y at strictly increasing arguments J."""
    J = np.asarray(J, np.float64)
    y = np.asarray(y, np.float64)
    j = np.asarray(j, np.float64)
    k = spline_interval(J, j)
    f = (j - J[k]) / (J[k + 1] - J[k])
    return alpha * j + beta + (1 - f) * y[k] + f * y[k + 1]


def spline_loss_grad(J, y, stride, alpha, beta, pos):
    """Mean squared residual of the model on samples pos[t] = select((t+1)*stride)
    (those outside [J[0], J[-1]] unseen, the mean still over all of them) and its
    gradient with respect to the knots (float64, closed form)."""
    J = np.asarray(J, np.float64)
    y = np.asarray(y, np.float64)
    pos = np.asarray(pos, np.float64)
    ns = pos.size
    j = (np.arange(ns, dtype=np.float64) + 1) * stride
    keep = (j >= J[0]) & (j <= J[-1])
    j, pos = j[keep], pos[keep]
    k = spline_interval(J, j)
    f = (j - J[k]) / (J[k + 1] - J[k])
    e = alpha * j + beta + (1 - f) * y[k] + f * y[k + 1] - pos
    loss = float((e * e).sum() / ns)
    g = np.zeros_like(y)
    np.add.at(g, k, 2.0 / ns * e * (1 - f))
    np.add.at(g, k + 1, 2.0 / ns * e * f)
    return loss, g


def place_knots(ones_pos, tol, gap):
    """Greedy error-cone knot placement of the learned select (pure Python).

    ones_pos[i] = position of the (i+1)-th one.  Arguments are cut into chunks
    of `gap`; in each chunk a knot sits on the first one, and from the last knot
    (x0, y0) the cone of slopes s with |y0 + s*(j - x0) - pos(j)| <= tol for every
    one seen shrinks until it is empty; the knot then goes to the previous
    argument, on the cone's middle slope.  The chunk's last argument closes it.
    Returns (J, value): knot arguments and float64 knot positions (every one is
    within tol of the interpolation).  Float operations in the order of the GPU
    placement kernel, unfused."""
    m = len(ones_pos)
    J, V = [], []
    for c in range((m + gap - 1) // gap):
        j0 = c * gap + 1
        j1 = min((c + 1) * gap, m)
        x0 = float(j0)
        y0 = float(ones_pos[j0 - 1])
        lo, hi = -1e300, 1e300
        jp = j0
        J.append(j0)
        V.append(y0)
        for jj in range(j0 + 1, j1 + 1):
            P = float(ones_pos[jj - 1])
            dx = float(jj) - x0
            nlo = max(lo, (P - tol - y0) / dx)
            nhi = min(hi, (P + tol - y0) / dx)
            if nlo > nhi:
                ye = y0 + 0.5 * (lo + hi) * (float(jp) - x0)
                J.append(jp)
                V.append(ye)
                x0, y0 = float(jp), ye
                dx = float(jj) - x0
                nlo = (P - tol - y0) / dx
                nhi = (P + tol - y0) / dx
            lo, hi = nlo, nhi
            jp = jj
        if float(jp) > x0:
            J.append(jp)
            V.append(y0 + 0.5 * (lo + hi) * (float(jp) - x0))
    return np.array(J, np.uint64), np.array(V, np.float64)


def model_line(alpha, beta, j):
    """alpha*j + beta, unfused (as the GPU model computes it)."""
    return alpha * np.asarray(j, np.float64) + beta
