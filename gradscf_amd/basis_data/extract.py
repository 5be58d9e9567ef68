"""Regenerate the basis JSON files of this directory from the reference's
vendored PySCF basis snapshot (NWChem-format .dat files).

Parsing follows the reference's snapshot reader
(src/gradscf/integrals/basis/data/__init__.py:152-205): an element's block is
found between "BASIS SET" / "END" delimiters; every shell header starts a
new shell of its angular momentum, an SP header starts an s and a p shell fed
from the first and second coefficient columns; "D" exponents become "e";
the element's shells are then ordered s, then p, then d (file order inside
each angular momentum).  Named sets resolve to files as the reference's
loader does (Pople names through its fixed alias table,
`__init__.py:22-60`; the others by normalised file stem).

Output: {element: [[l, [exponent, c_1, ..., c_nctr], ...], ...]} written with
json.dumps(indent=0), for the elements H, He, C, N, O.

Usage: python gradscf_amd/basis_data/extract.py [SNAPSHOT_DIR]
(default SNAPSHOT_DIR: the reference checkout's
src/gradscf/integrals/basis/data/pyscf_basis_snapshot)
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
DEFAULT_SNAPSHOT = Path("/root/reference/src/gradscf/integrals/basis/data/pyscf_basis_snapshot")
ELEMENTS = ("H", "He", "C", "N", "O")
L_OF = {"S": 0, "P": 1, "D": 2, "F": 3, "G": 4, "H": 5, "I": 6}

# output file stem -> snapshot file (relative to the snapshot root)
SETS = {
    "sto-3g": "sto-3g.dat",
    "6-31g": "pople-basis/6-31G.dat",
    "6-31gs": "pople-basis/6-31Gs.dat",
    "cc-pvdz": "cc-pvdz.dat",
    # auxiliary (density-fitting) sets
    "cc-pvdz-jkfit": "cc-pvdz-jkfit.dat",
    "def2-universal-jkfit": "def2-universal-jkfit.dat",
}

_DELIM = re.compile(r"#\s*BASIS SET.*\n|END\n")


def element_lines(text: str, symbol: str) -> list[str]:
    """The lines of `symbol`'s block: a chunk whose first token is the symbol,
    or, in a chunk that opens with comments, the lines from the first data
    line on when that line names the symbol."""
    for chunk in _DELIM.split(text):
        head = chunk.split(None, 1)
        if not head:
            continue
        if head[0] == symbol:
            return chunk.splitlines()
        if head[0].startswith("#"):
            lines = chunk.splitlines()
            for i, line in enumerate(lines):
                s = line.strip()
                if not s or s.startswith("#"):
                    continue
                if line.split(None, 1)[0] == symbol:
                    return lines[i:]
                break
    raise KeyError(f"no block for {symbol!r}")


def parse_block(lines: list[str]) -> list[list]:
    by_l: dict[int, list[list]] = {}
    label = None
    for raw in lines:
        line = raw.split("#")[0].strip()
        up = line.upper()
        if not line or up.startswith("END") or up.startswith("BASIS"):
            continue
        if line[0].isalpha():
            words = line.split()
            label = (words[0] if len(words) == 1 else words[1]).upper()
            if label == "SP":
                by_l.setdefault(0, []).append([0])
                by_l.setdefault(1, []).append([1])
            elif label in L_OF:
                by_l.setdefault(L_OF[label], []).append([L_OF[label]])
            else:
                raise ValueError(f"shell label {label!r} not supported")
            continue
        if label is None:
            raise ValueError("primitive line before a shell header")
        values = [float(w.replace("D", "e").replace("d", "e")) for w in line.split()]
        if label == "SP":
            by_l[0][-1].append([values[0], values[1]])
            by_l[1][-1].append([values[0], values[2]])
        else:
            by_l[L_OF[label]][-1].append(values)
    shells = [shell for l in sorted(by_l) for shell in by_l[l]]
    if not shells:
        raise ValueError("empty basis block")
    return shells


def extract(snapshot: Path, rel: str, elements=ELEMENTS) -> dict:
    text = (snapshot / rel).read_text(encoding="utf-8")
    out = {}
    for sym in elements:
        try:
            out[sym] = parse_block(element_lines(text, sym))
        except KeyError:
            continue
    return out


def main(argv) -> None:
    snapshot = Path(argv[1]) if len(argv) > 1 else DEFAULT_SNAPSHOT
    for stem, rel in SETS.items():
        data = extract(snapshot, rel)
        (HERE / f"{stem}.json").write_text(json.dumps(data, indent=0))
        print(f"{stem}.json: {', '.join(data)}")


if __name__ == "__main__":
    main(sys.argv)
