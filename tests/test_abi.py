"""The C-ABI library loads and exports every entry point its headers declare
(no compute call: this runs without a GPU)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from topaz_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text):
            name = m.group(1)
            if name in ("if", "for", "while", "return", "sizeof"):
                continue
            names.append((h.name, name))
    return names


def test_headers_declare_something():
    assert len(declared_functions()) >= 8


@pytest.mark.parametrize("header,name", declared_functions())
def test_symbol_exported(header, name):
    L = C.CDLL(str(_lib.LIB_PATH))
    assert hasattr(L, name), "%s declares %s but the library does not export it" % (header, name)


def test_version_string():
    assert b"gfx950" in _lib.lib().tpz_version()
