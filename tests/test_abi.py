"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU needed: nothing is called except mcx_version / mcx_last_error)."""
import ctypes as C
import glob
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(mcx_\w+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    from montecarlox import _lib
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (mcx_\w+)", out))
    assert names <= exported


def test_binding_covers_header():
    import montecarlox  # noqa: F401  (registers every module's signatures)
    from montecarlox import _lib
    assert declared() == set(_lib.exported_symbols())


def test_version_and_error_calls_without_gpu():
    from montecarlox import _lib
    lib = _lib.load()
    a, b = C.c_int32(), C.c_int32()
    assert lib.mcx_version(C.byref(a), C.byref(b)) == 0
    assert (a.value, b.value) >= (0, 1)
    assert isinstance(lib.mcx_last_error(), bytes)


def test_product_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "montecarlox.jl_amd")
    for path in glob.glob(os.path.join(pkg, "**", "*.*"), recursive=True):
        if path.endswith((".py", ".cpp", ".hip", ".h")):
            txt = open(path).read()
            assert "import oracle" not in txt and "from oracle" not in txt, path
            assert "mcx_oracle" not in txt and "liboracle" not in txt, path
