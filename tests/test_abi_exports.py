"""The C-ABI library loads and exports every function include/*.h declares
(no GPU needed: nothing is called)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _declared(prefix="hg_"):
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(" + prefix + r"\w+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    from libhybridgrains_amd import _lib

    L = _lib.lib()
    names = _declared()
    assert len(names) > 10
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_version_string():
    from libhybridgrains_amd import _lib

    assert _lib.lib().hg_version().decode().startswith("hg ")


def test_host_library_exports_scene_readers():
    from libhybridgrains_amd import xmlscene

    L = xmlscene.host_lib()
    names = _declared("hgx_")
    assert len(names) >= 15
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
