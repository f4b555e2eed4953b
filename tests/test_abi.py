"""The C-ABI library loads (no GPU needed) and exports every entry point
include/gcr.h declares; the Python binding's export list matches the header.
No compute call is made here."""
import ctypes
import os
import re

from pygcransac import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "gcr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gcr_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("gcr_find_homography", "gcr_find_fundamental_matrix", "gcr_find_essential_matrix",
                 "gcr_find_batch", "gcr_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.library_path())
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_export_list_matches_header():
    assert sorted(_native.EXPORTS) == _declared()
