"""The C-ABI library loads and exports every function include/*.h declares
(no device calls; runs without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        for m in re.finditer(r"\b(psx_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "psx_mcm" in names and "psx_cov" in names and "psx_decouple_covmat" in names
    assert len(names) >= 20


def test_library_exports_all_declared(pkg):
    from importlib import import_module
    abi = import_module(pkg.__name__ + "._abi")
    lib = ctypes.CDLL(abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers every declared entry point
    assert set(declared_functions()) <= set(abi.SIGNATURES)


def test_cov_arity_host_only(pkg):
    from importlib import import_module
    abi = import_module(pkg.__name__ + "._abi")
    assert abi.cov_arity(0) == (8, 4, 4)
    assert abi.cov_arity(3) == (5, 4, 2)
    assert abi.cov_arity(6) == (2, 4, 0)


def test_request_struct_layout_matches_header(pkg, tmp_path):
    """ctypes' psx_request has the C layout of include/powerspectra_hip.h."""
    import shutil
    import subprocess
    from importlib import import_module
    gcc = shutil.which("gcc") or shutil.which("cc")
    if gcc is None:
        import pytest
        pytest.skip("no C compiler")
    abi = import_module(pkg.__name__ + "._abi")
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "powerspectra_hip.h"\n'
        'int main(void) { printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(psx_request), '
        'offsetof(psx_request, LW), offsetof(psx_request, W), offsetof(psx_request, spectra), '
        'offsetof(psx_request, ratios), offsetof(psx_request, out), offsetof(psx_request, ld)); '
        'return 0; }\n')
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    R = abi.PsxRequest
    want = [ctypes.sizeof(R), R.LW.offset, R.W.offset, R.spectra.offset, R.ratios.offset,
            R.out.offset, R.ld.offset]
    assert got == want
