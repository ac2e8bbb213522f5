"""The C-ABI library loads and exports every symbol the public headers
declare (no compute calls)."""
import re
from pathlib import Path

import crushrs_amd
from crushrs_amd import _abi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(crush_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python mirror binds each of them
    assert syms <= set(_abi.SIGNATURES), sorted(syms - set(_abi.SIGNATURES))


def test_abi_version_and_error_string():
    lib = _abi.load()
    assert lib.crush_abi_version() == 3
    assert isinstance(lib.crush_last_error(), bytes)
    assert lib.crush_device_count() >= 0
