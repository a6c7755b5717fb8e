"""Regenerate the substitution-matrix tables of topaz_amd from the reference.

The eight scoring matrices (BLOSUM45/50/62/80/90, PAM30/70/250) are published
scoring data that the reference ships as 25x25 int8 arrays in the residue order
ARNDCQEGHILKMFPSTWYVBJZX* (src/substitution.h:20-258).  Scores must be
bit-identical, so the values are carried over unchanged; this script re-emits
them in this package's own formats:

  topaz_amd/csrc/matrices.inc   C initialisers used by the native library
  topaz_amd/data/matrices.json  the same values for the Python host layer

Usage (only where the reference checkout exists):
    python topaz_amd/data/gen_matrices.py /root/reference/src/substitution.h
"""
import json
import re
import sys
from pathlib import Path

NAMES = ["BLOSUM45", "BLOSUM50", "BLOSUM62", "BLOSUM80", "BLOSUM90",
         "PAM30", "PAM70", "PAM250"]
ORDER = "ARNDCQEGHILKMFPSTWYVBJZX*"


def parse(header_text):
    out = {}
    for name in NAMES:
        m = re.search(r"\b%s\s*\[\s*625\s*\]\s*=\s*\{(.*?)\};" % name,
                      header_text, re.S)
        if not m:
            raise SystemExit("matrix %s not found" % name)
        body = re.sub(r"//[^\n]*", "", m.group(1))
        vals = [int(v) for v in re.findall(r"-?\d+", body)]
        if len(vals) != 625:
            raise SystemExit("matrix %s has %d values" % (name, len(vals)))
        out[name] = vals
    return out


def main():
    src = Path(sys.argv[1] if len(sys.argv) > 1
               else "/root/reference/src/substitution.h")
    mats = parse(src.read_text())
    here = Path(__file__).resolve().parent
    (here / "matrices.json").write_text(
        json.dumps({"order": ORDER, "matrices": mats}, indent=0) + "\n")
    lines = ["/* Generated by topaz_amd/data/gen_matrices.py: 25x25 scores,",
             " * W[s * 25 + q], residue order %s. */" % ORDER]
    for name in NAMES:
        v = mats[name]
        lines.append("static const int8_t kMat_%s[625] = {" % name)
        for r in range(25):
            lines.append("  " + ",".join("%3d" % x for x in v[25 * r:25 * r + 25]) + ",")
        lines.append("};")
    lines.append("static const struct { const char* name; const int8_t* m; } kMatrices[] = {")
    for name in NAMES:
        lines.append('  {"%s", kMat_%s},' % (name, name))
    lines.append("};")
    (here.parent / "csrc" / "matrices.inc").write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
