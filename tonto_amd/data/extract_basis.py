"""Extract the basis sets this package needs from Tonto's basis library into JSON.

Tonto keeps its library as keyed text files (`basis_sets/<name>`) in two layouts:
``keys= { gamess-us= }`` (shell header ``S   3`` followed by ``i  exponent  coeff``
lines) and ``keys= { turbomole= }`` (header ``3  s`` followed by ``exponent  coeff``).
The reader semantics follow `foofiles/basis.foo` / `foofiles/shell.foo` (coefficients
are given for *normalised* primitives; Tonto un-normalises them after reading,
`foofiles/molecule.main.foo:803-807`).

Run in a checkout that has the reference library mounted:

    python tonto_amd/data/extract_basis.py /root/reference/basis_sets

It writes `tonto_amd/data/basis_sets.json` (exponents and coefficients exactly as
printed in the library, for H, C, N and O).  The JSON file is committed; nothing at
run time reads the reference tree.
"""
import json
import os
import re
import sys

BASES = ["STO-3G", "6-31G(d)", "6-31G(d,p)", "6-311G(d,p)", "DZP",
         "cc-pVDZ", "aug-cc-pVDZ", "cc-pVTZ"]
ELEMENTS = ["H", "C", "N", "O"]
L_OF = {"s": 0, "p": 1, "d": 2, "f": 3, "g": 4}


def _strip_comment(line):
    i = line.find("!")
    return line if i < 0 else line[:i]


def parse_library(text):
    """Return {label: [(l, [exp..], [coef..]), ...]} for every labelled block."""
    lines = [_strip_comment(l).rstrip() for l in text.splitlines()]
    fmt = "turbomole" if "turbomole=" in text else "gamess-us"
    blocks = {}
    label_re = re.compile(r"^\s*([A-Z][a-z]?:[^\s{]+)\s*(\{)?\s*$")
    i = 0
    n = len(lines)
    while i < n:
        m = label_re.match(lines[i])
        if not m:
            i += 1
            continue
        label = m.group(1)
        i += 1
        if not m.group(2):                    # opening brace on a later line
            while i < n and lines[i].strip() != "{":
                i += 1
            i += 1
        shells = []
        while i < n and lines[i].strip() != "}":
            tok = lines[i].split()
            if not tok:
                i += 1
                continue
            if fmt == "gamess-us":
                kind, nprim = tok[0].lower(), int(tok[1])
            else:
                nprim, kind = int(tok[0]), tok[1].lower()
            exps, coefs = [], []
            for k in range(nprim):
                t = lines[i + 1 + k].split()
                if fmt == "gamess-us":
                    t = t[1:]
                exps.append(float(t[0].replace("D", "E")))
                coefs.append(float(t[1].replace("D", "E")))
            shells.append((L_OF[kind], exps, coefs))
            i += 1 + nprim
        blocks[label] = shells
        i += 1
    return blocks


def main(libdir, out):
    data = {}
    for b in BASES:
        with open(os.path.join(libdir, b)) as f:
            blocks = parse_library(f.read())
        data[b] = {}
        for el in ELEMENTS:
            key = f"{el}:{b}"
            if key not in blocks:
                raise SystemExit(f"{key} missing from {b}")
            data[b][el] = [[l, e, c] for (l, e, c) in blocks[key]]
    with open(out, "w") as f:
        json.dump(data, f, indent=0)
    print(f"wrote {out}")


if __name__ == "__main__":
    lib = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/basis_sets"
    main(lib, os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis_sets.json"))
