"""Build recipe of the engine library aran_amd/libaran_amd.so (gfx950).

Every source is compiled with hipcc for gfx950 only; packed FP32 VALU
instructions are disabled for device code.  Objects go to aran_amd/_build/,
the shared library next to this file.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libaran_amd.so")
# (source, object name, extra flags); the tiled M2L is compiled once per
# range of expansion orders so its template instances build in parallel.
M2L_RANGES = [(1, 6), (7, 9), (10, 11), (12, 13), (14, 15)]
SOURCES = [("tree.hip", "tree", []), ("operators.hip", "operators", []),
           ("m2l_prep.hip", "m2l_prep", []), ("p2p.hip", "p2p", []),
           ("translate.hip", "translate", []),
           ("flavours.hip", "flavours", []),
           ("m2l_f64.hip", "m2l_f64", []),
           ("engine.hip", "engine", []), ("capi.cpp", "capi", []),
           ("server.cpp", "server", [])] + [
    ("m2l_tile.hip", "m2l_tile_%d" % a,
     ["-DARAN_M2L_PMIN=%d" % a, "-DARAN_M2L_PMAX=%d" % b])
    for a, b in M2L_RANGES] + [
    ("m2l_mfma.hip", "m2l_mfma_%d" % a,
     ["-DARAN_MFMA_PMIN=%d" % a, "-DARAN_MFMA_PMAX=%d" % b])
    for a, b in M2L_RANGES]
ARCH = "gfx950"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def _compile(item):
    src, name, extra = item
    obj = os.path.join(BUILD, name + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = (["hipcc", "-c"] + lang + [os.path.join(CSRC, src), "-o", obj] +
           COMMON + extra)
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(12, os.cpu_count() or 1)) as ex:
        for src, obj, r in ex.map(_compile, SOURCES):
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise RuntimeError("hipcc failed on " + src)
            if verbose and (r.stderr.strip()):
                sys.stderr.write(r.stderr)
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = ["hipcc", "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    build_driver()
    return LIB


DRIVER_SRC = os.path.join(ROOT, "tests", "c", "newton3_driver.c")
DRIVER = os.path.join(HERE, "newton3_driver")


def build_driver():
    """The C test driver (tests/c/newton3_driver.c): plain C against
    include/, linked to the engine library (found next to it at run time)."""
    cmd = ["cc", "-std=gnu99", "-O2", "-Wall", "-I" + os.path.join(ROOT,
                                                                  "include"),
           DRIVER_SRC, "-o", DRIVER + ".tmp", "-L" + HERE, "-laran_amd",
           "-Wl,-rpath,$ORIGIN", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("cc failed on " + DRIVER_SRC)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
