"""Python front end of libwtg.so (the gfx950 wavelet-tree / rank / select C ABI).

This is plumbing for tests and benchmarks: every call goes straight through
ctypes to the C ABI declared in include/wtg.h.  There is no CPU fallback: if
libwtg.so is missing or no gfx950 device is present the calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwtg.so")
NONE = (1 << 64) - 1

_lib = None

vp, u64, u32, i32 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
U64P = C.POINTER(C.c_uint64)


class WtgError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__("wtg error %d: %s" % (rc, msg))
        self.rc = rc


class Node(C.Structure):
    _fields_ = [("bitmapPtr", u64), ("length", u64), ("parent", u64),
                ("leftChild", u64), ("rightChild", u64)]


class Select(C.Structure):
    _fields_ = [("nbits", u64), ("arg_cnt", u64), ("sb", u64), ("n_long", u64),
                ("logn", u32), ("logn4", u32), ("bit", C.c_int32), ("_pad", C.c_int32),
                ("superblock", vp), ("mini", vp), ("mini_width", vp), ("long_index", vp),
                ("long_width", vp), ("long_pool", vp)]


class HuffShape(C.Structure):
    _fields_ = [("n", u64), ("bv_size", u64), ("sigma", u32), ("n_nodes", u32),
                ("n_inner", u32), ("max_depth", u32), ("C", u64 * 257),
                ("bv_pos", u64 * 511), ("bv_pos_rank", u64 * 511),
                ("child", (C.c_int32 * 2) * 511), ("parent", C.c_int32 * 511),
                ("sym", C.c_int32 * 511), ("c_to_leaf", C.c_int32 * 256),
                ("path", u64 * 256), ("path_len", u32 * 256)]


class LselParams(C.Structure):
    _fields_ = [("train_stride", u32), ("max_gap", u32), ("tol", u32),
                ("epochs", u32), ("lr", C.c_float), ("eps_max", u32)]


class LselInfo(C.Structure):
    _fields_ = [("nbits", u64), ("ones", u64), ("learned", u32), ("nknots", u32),
                ("eps", u64), ("eps_init", u64), ("loss_init", C.c_double),
                ("loss_final", C.c_double), ("trained", u32), ("lut_log", u32)]


class LselModelView(C.Structure):
    _fields_ = [("d_knots", C.c_void_p), ("nknots", u32), ("lut_log", u32),
                ("d_lut", C.c_void_p), ("lut_size", u64), ("alpha", C.c_double),
                ("beta", C.c_double), ("learned", u32), ("_pad", u32)]


# a knot of the learned-select model: {uint32 j; float y}
KNOT_DTYPE = np.dtype([("j", "<u4"), ("y", "<f4")])


# name -> (restype, argtypes); every symbol here is declared in include/wtg.h
SIGNATURES = {
    "wtg_version": (C.c_char_p, []),
    "wtg_last_error": (C.c_char_p, []),
    "wtg_device_count": (i32, [C.POINTER(i32)]),
    "wtg_ctx_create": (i32, [i32, C.POINTER(vp)]),
    "wtg_ctx_destroy": (i32, [vp]),
    "wtg_ctx_sync": (i32, [vp]),
    "wtg_ctx_set_blocking": (i32, [vp, i32]),
    "wtg_ctx_set_level_method": (i32, [vp, i32]),
    "wtg_ctx_stream": (vp, [vp]),
    "wtg_malloc": (i32, [vp, C.POINTER(vp), u64]),
    "wtg_free": (i32, [vp, vp]),
    "wtg_memcpy_h2d": (i32, [vp, vp, vp, u64]),
    "wtg_memcpy_d2h": (i32, [vp, vp, vp, u64]),
    "wtg_memcpy_d2d": (i32, [vp, vp, vp, u64]),
    "wtg_memset_zero": (i32, [vp, vp, u64]),
    "wtg_timer_start": (i32, [vp]),
    "wtg_timer_stop": (i32, [vp, C.POINTER(C.c_float)]),
    "wtg_kernel_timing": (i32, [vp, i32]),
    "wtg_kernel_time": (i32, [vp, C.c_char_p, C.POINTER(C.c_double), U64P]),
    "wtg_kernel_timing_reset": (i32, [vp]),
    "wtg_kernel_timing_only": (i32, [vp, C.c_char_p]),
    "wtg_levels_words": (u64, [u64, u32]),
    "wtg_levels_build": (i32, [vp, vp, i32, u64, u32, vp, vp, u64]),
    "wtg_remap": (i32, [vp, vp, i32, u64, U64P]),
    "wtg_wt_host": (i32, [vp, vp, i32, u64, u64, vp, vp]),
    "wtg_rank9_words": (u64, [u64]),
    "wtg_rank9_build": (i32, [vp, vp, u64, vp]),
    "wtg_rank9_batch": (i32, [vp, vp, vp, u64, vp, u64, vp]),
    "wtg_rank9_build_noalloc": (i32, [vp, vp, u64, vp]),
    "wtg_rank9_batch_bit": (i32, [vp, vp, vp, u64, i32, vp, u64, vp]),
    "wtg_select_build": (i32, [vp, vp, vp, u64, i32, C.POINTER(Select)]),
    "wtg_rank9_select_build": (i32, [vp, vp, u64, vp, C.POINTER(Select), C.POINTER(Select)]),
    "wtg_select_free": (i32, [vp, C.POINTER(Select)]),
    "wtg_pattern_words": (i32, [vp, vp, u64, u32, i32, vp]),
    "wtg_select_batch": (i32, [vp, C.POINTER(Select), vp, vp, u64, vp]),
    "wtg_wtint_build": (i32, [vp, vp, i32, u64, u32, C.POINTER(vp)]),
    "wtg_wtint_destroy": (i32, [vp, vp]),
    "wtg_wtint_from_tree": (i32, [vp, vp, u64, u32, u64, C.POINTER(vp)]),
    "wtg_wtint_info": (i32, [vp, U64P, C.POINTER(u32), U64P, C.POINTER(vp), C.POINTER(vp)]),
    "wtg_wtint_select1": (C.POINTER(Select), [vp]),
    "wtg_wtint_select0": (C.POINTER(Select), [vp]),
    "wtg_wtint_rank_batch": (i32, [vp, vp, vp, vp, u64, vp]),
    "wtg_wtint_access_batch": (i32, [vp, vp, vp, u64, vp]),
    "wtg_wtint_select_batch": (i32, [vp, vp, vp, vp, u64, vp]),
    "wtg_wtint_inverse_select_batch": (i32, [vp, vp, vp, u64, vp, vp]),
    "wtg_wtint_lex_count_batch": (i32, [vp, vp, vp, vp, vp, u64, vp, vp, vp]),
    "wtg_wtint_lex_smaller_count_batch": (i32, [vp, vp, vp, vp, u64, vp, vp]),
    "wtg_wtint_range_count_batch": (i32, [vp, vp, vp, vp, vp, vp, u64, vp]),
    "wtg_huff_build": (i32, [vp, vp, u64, C.POINTER(vp)]),
    "wtg_huff_destroy": (i32, [vp, vp]),
    "wtg_huff_from_bits": (i32, [vp, vp, vp, u64, C.POINTER(vp)]),
    "wtg_huff_get_shape": (i32, [vp, C.POINTER(HuffShape)]),
    "wtg_huff_data": (i32, [vp, C.POINTER(vp), C.POINTER(vp)]),
    "wtg_huff_select1": (C.POINTER(Select), [vp]),
    "wtg_huff_select0": (C.POINTER(Select), [vp]),
    "wtg_huff_rank_batch": (i32, [vp, vp, vp, vp, u64, vp]),
    "wtg_huff_access_batch": (i32, [vp, vp, vp, u64, vp, vp]),
    "wtg_huff_select_batch": (i32, [vp, vp, vp, vp, u64, vp]),
    "wtg_suffix_array": (i32, [vp, u64, vp]),
    "wtg_suffix_array_gpu": (i32, [vp, vp, u64, vp]),
    "wtg_bwt_from_sa_gpu": (i32, [vp, vp, vp, u64, vp]),
    "wtg_bwt_from_sa": (i32, [vp, vp, u64, vp]),
    "wtg_fm_count_batch": (i32, [vp, vp, vp, u32, u64, vp]),
    "wtg_lsel_build": (i32, [vp, u32, C.POINTER(vp), U64P, C.POINTER(LselParams), C.POINTER(vp)]),
    "wtg_lsel_destroy": (i32, [vp, vp]),
    "wtg_lsel_info": (i32, [vp, u32, C.POINTER(LselInfo)]),
    "wtg_lsel_model": (i32, [vp, u32, C.POINTER(LselModelView)]),
    "wtg_lsel_select_batch": (i32, [vp, vp, vp, vp, u64, vp]),
    "wtg_lsel_plain_batch": (i32, [vp, vp, vp, vp, u64, vp]),
    "wtg_lsel_sectors": (i32, [vp, vp, vp, vp, vp, u64, i32, vp]),
    "wtg_lsel_loss_grad": (i32, [vp, vp, u32, u32, C.c_double, C.c_double, vp, u64,
                                 C.POINTER(C.c_double), vp]),
    "wtg_gen_select_vectors": (i32, [u64, u32, u64, vp, vp]),
    "wtg_gen_batch": (i32, [u64, u32, C.c_double, C.c_double, vp, u32, vp, vp, vp]),
    "wtg_batch_create": (i32, [vp, vp, u32, vp, vp, vp, C.POINTER(vp)]),
    "wtg_batch_create_planned": (i32, [vp, vp, u32, vp, vp, vp, vp, C.POINTER(vp)]),
    "wtg_batch_tune": (i32, [vp, vp, u32, vp, vp, vp, u32, vp, vp]),
    "wtg_batch_destroy": (i32, [vp, vp]),
    "wtg_batch_set_graph": (i32, [vp, i32]),
    "wtg_batch_run": (i32, [vp, vp]),
    "wtg_batch_words": (i32, [vp, C.POINTER(vp), U64P]),
    "wtg_batch_seq": (i32, [vp, u32, C.POINTER(u32), U64P, U64P]),
    "wtg_batch_extract": (i32, [vp, vp, u32, vp]),
    "wtg_splitmix64": (u64, [U64P]),
    "wtg_gen_uniform": (i32, [u64, u64, u64, i32, vp]),
    "wtg_gen_geometric": (i32, [u64, u64, C.c_double, u64, i32, vp]),
    "wtg_gen_zipf_bytes": (i32, [u64, u64, C.c_double, vp]),
    "wtg_gen_dna": (i32, [u64, u64, vp]),
    "wtg_gen_patterns": (i32, [u64, vp, u64, u32, u64, vp]),
}


def lib():
    """Load libwtg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libwtg.so not built: run wavelet-suffix-fm-index_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise WtgError(rc, lib().wtg_last_error().decode())
    return rc


def _p(a):
    return C.c_void_p(a.ctypes.data)


def device_count():
    n = C.c_int(0)
    _check(lib().wtg_device_count(C.byref(n)))
    return n.value


def levels_of_sigma(sigma):
    """L = max(1, log2Up(sigma))  (wavelet/levelWT/WT.C:53)."""
    sigma = int(sigma)
    return max(1, (sigma - 1).bit_length()) if sigma > 1 else 1


def select_content_sha256(c):
    """SHA-256 of a select_support_mcl content in one fixed byte order
    (select/select_support_mcl_par.hpp:483-561 fields): the header
    (arg_cnt, sb, n_long, logn, logn4 as little-endian uint64), then
    superblock[sb] (uint64), mini[sb*64] (uint32), mini_width[sb] (uint8),
    long_width[sb] (uint8), long_index[sb] (int64) and the long pool
    [n_long*4096] (uint64).  `c` is a dict with those keys (SelectHandle.content)."""
    import hashlib
    h = hashlib.sha256()
    sb, nl = int(c["sb"]), int(c["n_long"])
    h.update(np.array([c["arg_cnt"], sb, nl, c["logn"], c["logn4"]], "<u8").tobytes())
    for key, dt, cnt in (("superblock", "<u8", sb), ("mini", "<u4", sb * 64), ("mini_width", "u1", sb),
                         ("long_width", "u1", sb), ("long_index", "<i8", sb), ("long_pool", "<u8", nl * 4096)):
        a = np.ascontiguousarray(np.asarray(c[key])[:cnt], dtype=dt)
        assert a.size == cnt, key
        h.update(a.tobytes())
    return h.hexdigest()


# ------------------------------------------------------------ generators
def gen_uniform(seed, n, rng, dtype=np.uint32):
    out = np.empty(n, dtype=dtype)
    _check(lib().wtg_gen_uniform(seed, n, rng, out.itemsize, _p(out)))
    return out


def gen_geometric(seed, n, p, cap, dtype=np.uint32):
    out = np.empty(n, dtype=dtype)
    _check(lib().wtg_gen_geometric(seed, n, p, cap, out.itemsize, _p(out)))
    return out


def gen_zipf_bytes(seed, n, s=1.1):
    out = np.empty(n, dtype=np.uint8)
    _check(lib().wtg_gen_zipf_bytes(seed, n, s, _p(out)))
    return out


def gen_select_vectors(seed, nvec, nbits):
    """(words[nvec, ceil(nbits/64)], density[nvec]) of the learned-select workload."""
    nw = (nbits + 63) // 64
    w = np.empty((nvec, nw), np.uint64)
    d = np.empty(nvec, np.float64)
    _check(lib().wtg_gen_select_vectors(seed, nvec, nbits, _p(w), _p(d)))
    return w, d


def gen_batch(seed, nseq, log2_min=10.0, log2_max=22.0, alphabets=(4, 20, 256), symbols=True):
    """(n[nseq], sigma[nseq], packed symbols or None) of the batched workload."""
    n = np.empty(nseq, np.uint64)
    sg = np.empty(nseq, np.uint32)
    al = np.ascontiguousarray(alphabets, np.uint32)
    _check(lib().wtg_gen_batch(seed, nseq, log2_min, log2_max, _p(al), al.size, _p(n), _p(sg), None))
    sym = None
    if symbols:
        sym = np.empty(int(n.sum()), np.uint8)
        _check(lib().wtg_gen_batch(seed, nseq, log2_min, log2_max, _p(al), al.size, _p(n), _p(sg),
                                   _p(sym)))
    return n, sg, sym


def gen_dna(seed, n):
    """n bytes of order-5 Markov DNA with repeats and N's, text[n-1] = 0."""
    out = np.empty(n, dtype=np.uint8)
    _check(lib().wtg_gen_dna(seed, n, _p(out)))
    return out


def gen_patterns(seed, text, m, npat):
    """npat x m patterns: even rows exact substrings, odd rows 1-2 substitutions."""
    text = np.ascontiguousarray(text, dtype=np.uint8)
    out = np.empty((npat, m), dtype=np.uint8)
    _check(lib().wtg_gen_patterns(seed, _p(text), text.size, m, npat, _p(out)))
    return out


def suffix_array(text):
    """Suffix array of a byte text (host, SA-IS)."""
    text = np.ascontiguousarray(text, dtype=np.uint8)
    sa = np.empty(text.size, dtype=np.int64)
    _check(lib().wtg_suffix_array(_p(text), text.size, _p(sa)))
    return sa


def suffix_array_gpu(ctx, d_text, n):
    """Suffix array of a device byte text on the GPU (prefix doubling); returns
    the device array (int64)."""
    d_sa = ctx.empty(n, np.int64)
    try:
        _check(lib().wtg_suffix_array_gpu(ctx.h, d_text.ptr, n, d_sa.ptr))
    except Exception:
        d_sa.free()
        raise
    return d_sa


def bwt_from_sa_gpu(ctx, d_text, d_sa, n):
    """BWT from a device text and suffix array on the GPU; returns the device array."""
    d_bwt = ctx.empty(n, np.uint8)
    try:
        _check(lib().wtg_bwt_from_sa_gpu(ctx.h, d_text.ptr, d_sa.ptr, n, d_bwt.ptr))
    except Exception:
        d_bwt.free()
        raise
    return d_bwt


def bwt_from_sa(text, sa):
    text = np.ascontiguousarray(text, dtype=np.uint8)
    sa = np.ascontiguousarray(sa, dtype=np.int64)
    out = np.empty(text.size, dtype=np.uint8)
    _check(lib().wtg_bwt_from_sa(_p(text), _p(sa), text.size, _p(out)))
    return out


# ------------------------------------------------------------ device side
class DeviceArray:
    """A device allocation with a numpy dtype/shape (owned by a Context)."""

    def __init__(self, ctx, nbytes, dtype=np.uint64, shape=None):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = shape if shape is not None else (self.nbytes // self.dtype.itemsize,)
        p = C.c_void_p()
        _check(lib().wtg_malloc(ctx.h, C.byref(p), max(1, self.nbytes)))
        self.ptr = p

    @property
    def size(self):
        return int(np.prod(self.shape))

    def free(self):
        if self.ptr is not None and self.ptr.value:
            _check(lib().wtg_free(self.ctx.h, self.ptr))
        self.ptr = None

    def get(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            _check(lib().wtg_memcpy_d2h(self.ctx.h, _p(out), self.ptr, out.nbytes))
        return out

    def zero(self):
        _check(lib().wtg_memset_zero(self.ctx.h, self.ptr, self.nbytes))


class SelectHandle:
    """Owns a wtg_select built on the device."""

    def __init__(self, ctx, st, owned=True):
        self.ctx, self.st, self.owned = ctx, st, owned

    @property
    def arg_cnt(self):
        return int(self.st.arg_cnt)

    def content(self):
        """Copy the structure back: dict of numpy arrays (sizes exact)."""
        s, ctx = self.st, self.ctx
        sb, nl = int(s.sb), int(s.n_long)
        out = {"arg_cnt": int(s.arg_cnt), "sb": sb, "n_long": nl, "logn": int(s.logn),
               "logn4": int(s.logn4), "bit": int(s.bit)}

        def cp(ptr, n, dt):
            a = np.empty(n, dtype=dt)
            if n:
                _check(lib().wtg_memcpy_d2h(ctx.h, _p(a), C.c_void_p(ptr), a.nbytes))
            return a
        out["superblock"] = cp(s.superblock, sb, np.uint64)
        out["mini"] = cp(s.mini, sb * 64, np.uint32)
        out["mini_width"] = cp(s.mini_width, sb, np.uint8)
        out["long_index"] = cp(s.long_index, sb, np.int64)
        out["long_width"] = cp(s.long_width, sb, np.uint8)
        out["long_pool"] = cp(s.long_pool, nl * 4096, np.uint64) if nl else np.zeros(0, np.uint64)
        return out

    def free(self):
        if self.owned and self.st is not None:
            _check(lib().wtg_select_free(self.ctx.h, C.byref(self.st)))
            self.st = None


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().wtg_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h is not None:
            lib().wtg_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        _check(lib().wtg_ctx_sync(self.h))

    def set_blocking(self, b):
        _check(lib().wtg_ctx_set_blocking(self.h, 1 if b else 0))

    def set_level_method(self, method):
        """0: multi-level direct passes (default); 1: one split pass per level."""
        _check(lib().wtg_ctx_set_level_method(self.h, int(method)))

    # ---- memory
    def empty(self, n, dtype=np.uint64):
        dt = np.dtype(dtype)
        return DeviceArray(self, int(n) * dt.itemsize, dt, (int(n),))

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.nbytes, a.dtype, a.shape)
        if a.nbytes:
            _check(lib().wtg_memcpy_h2d(self.h, d.ptr, _p(a), a.nbytes))
        return d

    # ---- timing
    def timer_start(self):
        _check(lib().wtg_timer_start(self.h))

    def timer_stop(self):
        ms = C.c_float(0)
        _check(lib().wtg_timer_stop(self.h, C.byref(ms)))
        return ms.value

    def kernel_timing(self, enable):
        _check(lib().wtg_kernel_timing(self.h, 1 if enable else 0))

    def kernel_time(self, name):
        ms, n = C.c_double(0), C.c_uint64(0)
        _check(lib().wtg_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def kernel_timing_reset(self):
        _check(lib().wtg_kernel_timing_reset(self.h))

    def kernel_timing_only(self, prefix):
        """time only launches whose name starts with prefix (None: all)"""
        _check(lib().wtg_kernel_timing_only(self.h, prefix.encode() if prefix else None))

    # ---- (a) levels
    def levels_build(self, d_sym, n, L, d_words, d_nodes=None, slots=0):
        _check(lib().wtg_levels_build(self.h, d_sym.ptr, d_sym.dtype.itemsize, int(n), int(L),
                                      d_words.ptr, d_nodes.ptr if d_nodes is not None else None,
                                      int(slots)))

    def levels(self, s, L, with_nodes=False):
        """Host convenience: returns (words, nodes[slots,5] or None)."""
        s = np.ascontiguousarray(s)
        n = s.size
        d = self.to_device(s)
        nw = lib().wtg_levels_words(n, L)
        dw = self.empty(max(nw, 1), np.uint64)
        dn = None
        slots = (1 << L) if with_nodes else 0
        if with_nodes:
            dn = DeviceArray(self, slots * 40, np.uint64, (slots, 5))
        try:
            self.levels_build(d, n, L, dw, dn, slots)
            words = dw.get()[:nw]
            nodes = dn.get() if with_nodes else None
        finally:
            d.free(); dw.free()
            if dn is not None:
                dn.free()
        return words, nodes

    def remap(self, d_sym):
        sig = C.c_uint64(0)
        _check(lib().wtg_remap(self.h, d_sym.ptr, d_sym.dtype.itemsize, d_sym.size, C.byref(sig)))
        return sig.value

    def wt_host(self, s, sigma):
        """The WT() drop-in path: host in, host out (levels = max(1, log2Up(sigma)))."""
        s = np.ascontiguousarray(s)
        L = levels_of_sigma(sigma)
        nw = lib().wtg_levels_words(s.size, L)
        words = np.zeros(max(nw, 1), np.uint64)
        nodes = np.zeros((2 * int(sigma), 5), np.uint64)
        _check(lib().wtg_wt_host(self.h, _p(s), s.itemsize, s.size, int(sigma), _p(words), _p(nodes)))
        return words[:nw], nodes

    # ---- (b) rank
    def rank9_build(self, d_words, nbits):
        D = self.empty(lib().wtg_rank9_words(int(nbits)), np.uint64)
        _check(lib().wtg_rank9_build(self.h, d_words.ptr, int(nbits), D.ptr))
        return D

    def rank9_batch(self, d_words, D, nbits, d_idx, d_out=None):
        nq = d_idx.size
        out = d_out if d_out is not None else self.empty(nq, np.uint64)
        _check(lib().wtg_rank9_batch(self.h, d_words.ptr, D.ptr, int(nbits), d_idx.ptr, nq, out.ptr))
        return out

    # ---- (b'') two-bit patterns: argument words of `10` (pattern 10) or `01` (1)
    def pattern_words(self, d_words, nbits, pattern, clip):
        out = self.empty(max(1, (int(nbits) + 63) // 64), np.uint64)
        _check(lib().wtg_pattern_words(self.h, d_words.ptr, int(nbits), int(pattern), int(bool(clip)), out.ptr))
        return out

    # ---- (b') select
    def rank9_select_build(self, d_words, nbits):
        D = self.empty(lib().wtg_rank9_words(int(nbits)), np.uint64)
        s1, s0 = Select(), Select()
        _check(lib().wtg_rank9_select_build(self.h, d_words.ptr, int(nbits), D.ptr,
                                            C.byref(s1), C.byref(s0)))
        return D, SelectHandle(self, s1), SelectHandle(self, s0)

    def select_build(self, d_words, D, nbits, bit):
        s = Select()
        _check(lib().wtg_select_build(self.h, d_words.ptr, D.ptr, int(nbits), int(bit), C.byref(s)))
        return SelectHandle(self, s)

    def select_batch(self, sel, d_words, d_i, d_out=None):
        nq = d_i.size
        out = d_out if d_out is not None else self.empty(nq, np.uint64)
        _check(lib().wtg_select_batch(self.h, C.byref(sel.st), d_words.ptr, d_i.ptr, nq, out.ptr))
        return out

    # ---- wt_int
    def wtint_build(self, d_sym, n, max_level=0):
        h = C.c_void_p()
        _check(lib().wtg_wtint_build(self.h, d_sym.ptr, d_sym.dtype.itemsize, int(n),
                                     int(max_level), C.byref(h)))
        return WtInt(self, h)

    # ---- wt_huff / FM
    def huff_build(self, d_text, n=None):
        h = C.c_void_p()
        n = d_text.size if n is None else n
        _check(lib().wtg_huff_build(self.h, d_text.ptr, int(n), C.byref(h)))
        return WtHuff(self, h)


class LearnedSelect:
    """Learned select over device bitvectors (each with a zero padding word)."""

    def __init__(self, ctx, d_vecs, nbits, params=None):
        self.ctx = ctx
        self.nvec = len(d_vecs)
        ptrs = (vp * self.nvec)(*[d.ptr.value for d in d_vecs])
        nb = (C.c_uint64 * self.nvec)(*[int(x) for x in nbits])
        h = C.c_void_p()
        pp = C.byref(params) if params is not None else None
        _check(lib().wtg_lsel_build(ctx.h, self.nvec, ptrs, nb, pp, C.byref(h)))
        self.h = h

    def info(self, b):
        i = LselInfo()
        _check(lib().wtg_lsel_info(self.h, b, C.byref(i)))
        return {k: getattr(i, k) for k, _ in LselInfo._fields_}

    def model(self, b):
        """(knots as a KNOT_DTYPE array, lookup table, lut_log, alpha, beta) of vector b."""
        m = LselModelView()
        _check(lib().wtg_lsel_model(self.h, b, C.byref(m)))
        knots = np.zeros(m.nknots, KNOT_DTYPE)
        lut = np.zeros(m.lut_size, np.uint32)
        if m.nknots:
            _check(lib().wtg_memcpy_d2h(self.ctx.h, _p(knots), C.c_void_p(m.d_knots), knots.nbytes))
            _check(lib().wtg_memcpy_d2h(self.ctx.h, _p(lut), C.c_void_p(m.d_lut), lut.nbytes))
        return knots, lut, m.lut_log, m.alpha, m.beta

    def query(self, d_vec, d_j, d_out, learned=True):
        fn = lib().wtg_lsel_select_batch if learned else lib().wtg_lsel_plain_batch
        _check(fn(self.ctx.h, self.h, d_vec.ptr, d_j.ptr, d_j.size, d_out.ptr))
        return d_out

    def sectors(self, d_vec, d_j, d_ans, learned=True):
        """Device uint32 array: 64-byte HBM sectors each query reads on the learned
        or the plain path (d_ans: that path's answers to the same queries)."""
        out = self.ctx.empty(d_j.size, np.uint32)
        _check(lib().wtg_lsel_sectors(self.ctx.h, self.h, d_vec.ptr, d_j.ptr, d_ans.ptr, d_j.size,
                                      1 if learned else 0, out.ptr))
        return out

    def select(self, vec, j, learned=True):
        ctx = self.ctx
        dv = ctx.to_device(np.ascontiguousarray(vec, np.uint32))
        dj = ctx.to_device(np.ascontiguousarray(j, np.uint64))
        do = ctx.empty(dj.size, np.uint64)
        try:
            self.query(dv, dj, do, learned)
            return do.get()
        finally:
            dv.free(); dj.free(); do.free()

    def free(self):
        if self.h is not None:
            _check(lib().wtg_lsel_destroy(self.ctx.h, self.h))
            self.h = None


def lsel_loss_grad(ctx, J, y, stride, alpha, beta, pos):
    """(loss, grad) of one spline model (knots at arguments J, values y) by the
    GPU training kernels."""
    knots = np.zeros(len(y), KNOT_DTYPE)
    knots["j"] = J
    knots["y"] = y
    dk = ctx.to_device(knots)
    dp = ctx.to_device(np.ascontiguousarray(pos, np.uint64))
    dg = ctx.empty(len(y), np.float32)
    loss = C.c_double(0)
    try:
        _check(lib().wtg_lsel_loss_grad(ctx.h, dk.ptr, len(y), stride, alpha, beta, dp.ptr,
                                        len(pos), C.byref(loss), dg.ptr))
        return loss.value, dg.get()
    finally:
        dk.free(); dp.free(); dg.free()


class Batch:
    """Batched level-wise trees of byte sequences packed in one device array."""

    def __init__(self, ctx, d_sym, n, sigma, off=None, plan=None):
        """plan: None (default groups) or 9 cap_log values (see tune())."""
        self.ctx = ctx
        n, sigma, off = Batch._arrays(n, sigma, off)
        self.n, self.sigma, self.off = n, sigma, off
        h = C.c_void_p()
        if plan is None:
            _check(lib().wtg_batch_create(ctx.h, d_sym.ptr, n.size, _p(off), _p(n), _p(sigma), C.byref(h)))
        else:
            p = np.ascontiguousarray(plan, np.uint8)
            _check(lib().wtg_batch_create_planned(ctx.h, d_sym.ptr, n.size, _p(off), _p(n), _p(sigma), _p(p),
                                                  C.byref(h)))
        self.h = h

    @staticmethod
    def _arrays(n, sigma, off):
        n = np.ascontiguousarray(n, np.uint64)
        sigma = np.ascontiguousarray(sigma, np.uint32)
        if off is None:
            off = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.uint64)
        return n, sigma, np.ascontiguousarray(off, np.uint64)

    @staticmethod
    def tune(ctx, d_sym, n, sigma, off=None, reps=2):
        """The autotuner's plan (9 cap_log values, 255 = default) and its timings
        (ms[L][c] of candidate c for alphabet width L, -1 = not tried)."""
        n, sigma, off = Batch._arrays(n, sigma, off)
        plan = np.zeros(9, np.uint8)
        ms = np.zeros(45, np.float32)
        _check(lib().wtg_batch_tune(ctx.h, d_sym.ptr, n.size, _p(off), _p(n), _p(sigma), reps, _p(plan), _p(ms)))
        return plan, ms.reshape(9, 5)

    def set_graph(self, on):
        _check(lib().wtg_batch_set_graph(self.h, 1 if on else 0))

    def run(self):
        _check(lib().wtg_batch_run(self.ctx.h, self.h))

    def seq(self, i):
        L, fb, st = C.c_uint32(), C.c_uint64(), C.c_uint64()
        _check(lib().wtg_batch_seq(self.h, i, C.byref(L), C.byref(fb), C.byref(st)))
        return L.value, fb.value, st.value

    def words(self):
        p, nw = C.c_void_p(), C.c_uint64()
        _check(lib().wtg_batch_words(self.h, C.byref(p), C.byref(nw)))
        a = np.empty(nw.value, np.uint64)
        _check(lib().wtg_memcpy_d2h(self.ctx.h, _p(a), p, a.nbytes))
        return a

    def extract(self, i):
        """Sequence i's tree bits in the WT() layout."""
        L = self.seq(i)[0]
        nw = lib().wtg_levels_words(int(self.n[i]), L)
        d = self.ctx.empty(max(nw, 1), np.uint64)
        try:
            _check(lib().wtg_batch_extract(self.ctx.h, self.h, i, d.ptr))
            return d.get()[:nw]
        finally:
            d.free()

    def free(self):
        if self.h is not None:
            _check(lib().wtg_batch_destroy(self.ctx.h, self.h))
            self.h = None


class WtHuff:
    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h
        self.shape = HuffShape()
        _check(lib().wtg_huff_get_shape(h, C.byref(self.shape)))
        w, d = C.c_void_p(), C.c_void_p()
        _check(lib().wtg_huff_data(h, C.byref(w), C.byref(d)))
        self.words_ptr, self.dir_ptr = w, d
        self.n, self.nbits = int(self.shape.n), int(self.shape.bv_size)

    def words(self):
        """Tree bits: ceil(bv_size/64) words and the zero padding word."""
        nw = (self.nbits + 63) // 64 + 1
        a = np.empty(nw, np.uint64)
        if self.words_ptr.value:
            _check(lib().wtg_memcpy_d2h(self.ctx.h, _p(a), self.words_ptr, a.nbytes))
        else:
            a[:] = 0
        return a

    def directory(self):
        nd = lib().wtg_rank9_words(self.nbits)
        a = np.empty(nd, np.uint64)
        _check(lib().wtg_memcpy_d2h(self.ctx.h, _p(a), self.dir_ptr, a.nbytes))
        return a

    def select1(self):
        return SelectHandle(self.ctx, lib().wtg_huff_select1(self.h).contents, owned=False)

    def select0(self):
        return SelectHandle(self.ctx, lib().wtg_huff_select0(self.h).contents, owned=False)

    def _q(self, fn, arrs, dts, n_out=1, null_out=False):
        ctx = self.ctx
        nq = arrs[0].size
        d_in = [ctx.to_device(np.ascontiguousarray(a, dtype=dt)) for a, dt in zip(arrs, dts)]
        d_out = [ctx.empty(nq, np.uint64) for _ in range(n_out)]
        try:
            outs = [d.ptr for d in d_out] + ([None] if null_out else [])
            _check(fn(ctx.h, self.h, *[d.ptr for d in d_in], nq, *outs))
            res = [d.get() for d in d_out]
        finally:
            for d in d_in + d_out:
                d.free()
        return res[0] if n_out == 1 else res

    def rank(self, i, c):
        return self._q(lib().wtg_huff_rank_batch, [i, c], [np.uint64, np.uint32])

    def access(self, i):
        return self._q(lib().wtg_huff_access_batch, [i], [np.uint64], null_out=True)

    def inverse_select(self, i):
        sym, rk = self._q(lib().wtg_huff_access_batch, [i], [np.uint64], n_out=2)
        return rk, sym

    def select(self, i, c):
        return self._q(lib().wtg_huff_select_batch, [i, c], [np.uint64, np.uint32])

    def fm_count(self, pats):
        """Backward-search counts; pats: (npat, m) uint8."""
        pats = np.ascontiguousarray(pats, dtype=np.uint8)
        npat, m = pats.shape
        ctx = self.ctx
        dp = ctx.to_device(pats.reshape(-1))
        do = ctx.empty(npat, np.uint64)
        try:
            _check(lib().wtg_fm_count_batch(ctx.h, self.h, dp.ptr, m, npat, do.ptr))
            return do.get()
        finally:
            dp.free(); do.free()

    def free(self):
        if self.h is not None:
            _check(lib().wtg_huff_destroy(self.ctx.h, self.h))
            self.h = None


class WtInt:
    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h
        n, L, sg = C.c_uint64(), C.c_uint32(), C.c_uint64()
        w, d = C.c_void_p(), C.c_void_p()
        _check(lib().wtg_wtint_info(h, C.byref(n), C.byref(L), C.byref(sg), C.byref(w), C.byref(d)))
        self.n, self.L, self.sigma = n.value, L.value, sg.value
        self.words_ptr, self.dir_ptr = w, d
        self.nbits = self.n * self.L

    def words(self):
        nw = (self.nbits + 63) // 64
        a = np.empty(nw, np.uint64)
        if nw:
            _check(lib().wtg_memcpy_d2h(self.ctx.h, _p(a), self.words_ptr, a.nbytes))
        return a

    def directory(self):
        nd = lib().wtg_rank9_words(self.nbits)
        a = np.empty(nd, np.uint64)
        _check(lib().wtg_memcpy_d2h(self.ctx.h, _p(a), self.dir_ptr, a.nbytes))
        return a

    def select1(self):
        return SelectHandle(self.ctx, lib().wtg_wtint_select1(self.h).contents, owned=False)

    def select0(self):
        return SelectHandle(self.ctx, lib().wtg_wtint_select0(self.h).contents, owned=False)

    def _q(self, fn, *arrs, n_out=1):
        ctx = self.ctx
        nq = arrs[0].size
        d_in = [ctx.to_device(np.ascontiguousarray(a, dtype=np.uint64)) for a in arrs]
        d_out = [ctx.empty(nq, np.uint64) for _ in range(n_out)]
        try:
            _check(fn(ctx.h, self.h, *[d.ptr for d in d_in], nq, *[d.ptr for d in d_out]))
            res = [d.get() for d in d_out]
        finally:
            for d in d_in + d_out:
                d.free()
        return res[0] if n_out == 1 else res

    def rank(self, i, c):
        return self._q(lib().wtg_wtint_rank_batch, i, c)

    def access(self, i):
        return self._q(lib().wtg_wtint_access_batch, i)

    def select(self, i, c):
        return self._q(lib().wtg_wtint_select_batch, i, c)

    def inverse_select(self, i):
        return self._q(lib().wtg_wtint_inverse_select_batch, i, n_out=2)

    def lex_count(self, i, j, c):
        """(rank(i,c), #values < c in [i,j), #values > c in [i,j)) per query."""
        return self._q(lib().wtg_wtint_lex_count_batch, i, j, c, n_out=3)

    def lex_smaller_count(self, i, c):
        """(rank(i,c), #values < c in [0,i)) per query."""
        return self._q(lib().wtg_wtint_lex_smaller_count_batch, i, c, n_out=2)

    def range_count(self, lb, rb, vlb, vrb):
        """#positions in [lb,rb] with values in [vlb,vrb] (range_search_2d count)."""
        return self._q(lib().wtg_wtint_range_count_batch, lb, rb, vlb, vrb)

    def free(self):
        if self.h is not None:
            _check(lib().wtg_wtint_destroy(self.ctx.h, self.h))
            self.h = None


from . import workloads  # noqa: E402,F401  (benchmark inputs and queries)
