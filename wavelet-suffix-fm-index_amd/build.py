"""Build libwtg.so (HIP kernels + C ABI) for gfx950, in-tree.

    python wavelet-suffix-fm-index_amd/build.py

Objects go to ``_build/`` next to this file; the shared library is
``libwtg.so`` in this directory.  hipcc cross-compiles for gfx950 without a GPU.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwtg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions anywhere (the path is integer-only anyway)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def _compile(args):
    cmd, out = args
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    hip, cpp = _sources()
    jobs_list = []
    for f in hip:
        o = os.path.join(OBJ, f + ".o")
        jobs_list.append(([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o], o))
    for f in cpp:
        o = os.path.join(OBJ, f + ".o")
        jobs_list.append((["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o], o))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, jobs_list))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB




def build_dropin_drivers(verbose=False):
    """g++ builds of tests/cpp/dropin_driver.cpp against the C++ drop-in headers
    (byte, -DINT and -DINT -DLONG symbol types), linked to libwtg.so."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_driver.cpp")
    outdir = os.path.join(ROOT, "tests", "cpp", "_build")
    os.makedirs(outdir, exist_ok=True)
    outs = []
    for name, defs in (("u8", []), ("u32", ["-DINT"]), ("u64", ["-DINT", "-DLONG"])):
        exe = os.path.join(outdir, "dropin_" + name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unused-variable"] + defs + [
            "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "host"), src,
            "-o", exe, "-L" + HERE, "-lwtg", "-Wl,-rpath," + HERE, "-Wl,-rpath,/opt/rocm/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("driver build failed: %s\n%s" % (" ".join(cmd), r.stderr))
        outs.append(exe)
    # serialize / load round trips of the C++ drop-ins
    exe = os.path.join(outdir, "serialize")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(HERE, "host"), os.path.join(ROOT, "tests", "cpp", "serialize_driver.cpp"),
           "-o", exe, "-L" + HERE, "-lwtg", "-Wl,-rpath," + HERE, "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed: %s\n%s" % (" ".join(cmd), r.stderr))
    outs.append(exe)
    if verbose:
        print("built", ", ".join(outs))
    return outs


if __name__ == "__main__":
    build(verbose=True)
    build_dropin_drivers(verbose=True)
    sys.exit(0)
