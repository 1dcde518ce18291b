"""The C-ABI library loads without a GPU and exports every function that
include/aran/*.h and include/vsg/*.h declare (no compute calls here)."""
import ctypes as C
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"typedef[^;]*;", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*", "", text)
        for m in re.finditer(r"\b((?:aran|vsg)_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from aran_amd import LIB_PATH
    lib = C.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_fails_loudly_when_missing(tmp_path, monkeypatch):
    import importlib

    import aran_amd._lib as L
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(L, "_lib", None)
    try:
        L.lib()
    except RuntimeError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("missing library was not reported")
    importlib.reload(L)


def test_headers_compile_as_c():
    """The public headers are plain C (gcc -fsyntax-only, no GLib)."""
    import subprocess
    src = "\n".join("#include <aran/%s>" % os.path.basename(h) for h in
                    sorted(glob.glob(os.path.join(ROOT, "include", "aran",
                                                  "*.h"))))
    r = subprocess.run(["gcc", "-fsyntax-only", "-std=c99", "-Wall",
                        "-I" + os.path.join(ROOT, "include"), "-x", "c", "-"],
                       input=src + "\nint main(void){return 0;}\n",
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
