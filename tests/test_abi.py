"""CPU checks of the C ABI boundary: libwtg.so loads, exports every function
declared in include/wtg.h (and the Python front end binds exactly those), and
the host-side workload generators are deterministic.  No GPU is needed."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "wtg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(wtg_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(wsfm):
    L = wsfm.lib()
    decl = declared_functions()
    assert len(decl) > 20
    missing = [n for n in decl if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(wsfm.SIGNATURES) == decl


def test_version_and_error_string(wsfm):
    assert wsfm.lib().wtg_version().startswith(b"wtg")
    assert isinstance(wsfm.lib().wtg_last_error(), bytes)


def test_size_helpers(wsfm):
    L = wsfm.lib()
    assert L.wtg_levels_words(100, 3) == 5
    assert L.wtg_levels_words(0, 8) == 0
    assert L.wtg_rank9_words(0) == 2
    assert L.wtg_rank9_words(512) == 4
    assert L.wtg_rank9_words(513) == 4
    assert L.wtg_rank9_words(4096) == 18


def test_generators_deterministic(wsfm):
    a = wsfm.gen_uniform(1, 100000, 1 << 20)
    b = wsfm.gen_uniform(1, 100000, 1 << 20)
    assert np.array_equal(a, b) and a.max() < (1 << 20)
    assert not np.array_equal(a, wsfm.gen_uniform(2, 100000, 1 << 20))
    g = wsfm.gen_geometric(3, 200000, 1e-4, (1 << 20) - 1)
    assert abs(g.mean() - (1 - 1e-4) / 1e-4) < 0.05 * 1e4
    z = wsfm.gen_zipf_bytes(4, 1 << 20)
    assert len(np.unique(z)) == 256
    c = np.sort(np.bincount(z, minlength=256))[::-1]
    assert c[0] > 20 * c[50]


def test_errors_without_device(wsfm):
    import ctypes
    h = ctypes.c_void_p()
    rc = wsfm.lib().wtg_ctx_create(-1, ctypes.byref(h))
    assert rc != 0
