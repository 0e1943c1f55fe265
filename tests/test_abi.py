"""The C-ABI library loads and exports every entry point include/rtpg.h declares
(no GPU work is issued here)."""
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "rtpg.h").read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(rtpg_[a-z0-9_]+)\(", text, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "rtpg_render" in names and "rtpg_sdtree_reset" in names and len(names) > 40


def test_library_exports_every_declared_symbol():
    from rtlib_amd import _lib

    lib = _lib.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes table covers the same set
    assert sorted(_lib.SIGNATURES) == _declared()


def test_version_string():
    from rtlib_amd import _lib

    assert b"gfx950" in _lib.load().rtpg_version()
