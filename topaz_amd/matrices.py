"""Substitution matrices (25 x 25, residue order ARNDCQEGHILKMFPSTWYVBJZX*),
the same values as the reference's substitution.h, and the alphabet maps of
lexicographical.h:19-74."""
import json
from pathlib import Path

import numpy as np

_DATA = json.loads((Path(__file__).resolve().parent / "data" / "matrices.json").read_text())
MATRIX_ORDER = _DATA["order"]
NAMES = tuple(_DATA["matrices"])

# The internal alphabet stores amino acid k of this ordering
# (lexicographical.h:7) as the letter chr(ord('A') + k).
INTERNAL_ORDER = "ACMLJIVTSKRQZEDBNHYFWXP*G"


def matrix(name="BLOSUM62"):
    """The named matrix as a flat int8 array of 625 entries, W[s * 25 + q]."""
    return np.array(_DATA["matrices"][name], dtype=np.int8)


def aa_table2():
    """Internal letter -> matrix index (lexicographical.h:19-33) as a 128-entry
    int8 table; bytes outside 'A'..'Y' / 'a'..'y' map to 23."""
    t = np.full(128, 23, dtype=np.int8)
    for k, aa in enumerate(INTERNAL_ORDER):
        idx = MATRIX_ORDER.index(aa)
        t[ord("A") + k] = idx
        t[ord("a") + k] = idx
    return t


def aa2internal():
    """Amino-acid letter -> internal letter (lexicographical.h:135-192)."""
    t = {}
    for k, aa in enumerate(INTERNAL_ORDER):
        t[aa] = chr(ord("A") + k)
        if aa.isalpha():
            t[aa.lower()] = chr(ord("a") + k)
    t["\n"] = "\n"
    return t
