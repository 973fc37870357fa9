"""The C-ABI library loads and exports every entry point include/*.h
declares (no GPU needed: nothing is called that touches a device)."""
import ctypes as C
import glob
import os
import re

import pytest

from hs_testlib import ROOT


def declared_symbols():
    syms = set()
    for hdr in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(hdr).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\**\s*(hs_\w+)\s*\(", text, flags=re.M):
            syms.add(m.group(1))
    return sorted(syms)


def test_header_declares_entries():
    syms = declared_symbols()
    assert "hs_cholesky_f64" in syms and "hs_cg_f64" in syms and "hs_symv_packed_f64" in syms
    assert len(syms) >= 30


@pytest.mark.parametrize("sym", declared_symbols())
def test_library_exports(hs, sym):
    lib = C.CDLL(hs.LIB_PATH)
    assert hasattr(lib, sym), "libhs_amd.so does not export %s" % sym


def test_binding_covers_header(hs):
    assert set(declared_symbols()) <= set(hs.exported_symbols())


def test_error_paths_without_device(hs):
    """Argument validation happens before any device work."""
    lib = hs.load()
    assert lib.hs_cholesky_f64(None, None, 10, 4) == -1
    assert lib.hs_handle_destroy(None) == -1
    assert lib.hs_sync(None) == -1
    assert lib.hs_malloc(None, 8) == -1
    assert lib.hs_last_error_string() is not None


def test_missing_library_fails_loudly(hs, tmp_path):
    with pytest.raises(RuntimeError):
        hs._lib_backup = hs._lib
        hs._lib = None
        try:
            hs.load(str(tmp_path / "nope.so"))
        finally:
            hs._lib = hs._lib_backup
