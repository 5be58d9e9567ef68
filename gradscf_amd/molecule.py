"""Molecule geometry: element table, unit conversion, nuclear repulsion, and the
deterministic geometry generators used by the J/K workloads.

Reference semantics followed:
  * Angstrom -> Bohr constant: src/gradscf/data/molecule.py:10
  * nuclear repulsion sum over i<j Z_i Z_j / r_ij: src/gradscf/data/molecule.py:186-198
  * atom-record parsing ("O 0 0 0; H ..." strings or (symbol, xyz) lists):
    src/gradscf/data/molecule.py:103-180
"""
from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np

ANGSTROM_TO_BOHR = 1.8897261245650618

_Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}


def atomic_number(symbol: str) -> int:
    clean = "".join(ch for ch in str(symbol) if ch.isalpha()).capitalize()
    if clean not in _Z:
        raise ValueError(f"Unsupported element symbol {symbol!r}.")
    return _Z[clean]


def std_symbol(symbol: str) -> str:
    clean = "".join(ch for ch in str(symbol) if ch.isalpha())
    if not clean:
        raise ValueError(f"Invalid atomic symbol {symbol!r}.")
    return clean[0].upper() + clean[1:].lower()


@dataclass(frozen=True)
class MoleculeSpec:
    """Atoms of one molecule; coordinates are stored in Bohr."""

    symbols: tuple
    coords_bohr: np.ndarray
    charge: int = 0
    spin: int = 0

    @property
    def charges(self) -> np.ndarray:
        return np.array([atomic_number(s) for s in self.symbols], dtype=np.int32)

    @property
    def natm(self) -> int:
        return len(self.symbols)

    @property
    def nelectron(self) -> int:
        return int(self.charges.sum()) - int(self.charge)

    def nuclear_repulsion(self) -> float:
        return nuclear_repulsion_energy(self)


def parse_molecule_spec(atom, *, unit="Angstrom", charge=0, spin=0) -> MoleculeSpec:
    """Accept "O 0 0 0; H 0 1 0" strings or [(symbol, (x, y, z)), ...] records."""
    records = []
    if isinstance(atom, str):
        for chunk in atom.replace("\n", ";").split(";"):
            parts = chunk.split()
            if not parts:
                continue
            if len(parts) != 4:
                raise ValueError(f"Invalid atom record {chunk!r}.")
            records.append((parts[0], tuple(float(v) for v in parts[1:])))
    else:
        for rec in atom:
            sym, xyz = rec[0], rec[1]
            records.append((sym, tuple(float(v) for v in xyz)))
    if not records:
        raise ValueError("Molecule has no atoms.")
    u = str(unit).lower()
    if u.startswith("a"):
        scale = ANGSTROM_TO_BOHR
    elif u.startswith("b"):
        scale = 1.0
    else:
        raise ValueError(f"Unknown unit {unit!r}.")
    symbols = tuple(std_symbol(s) for s, _ in records)
    for s in symbols:
        atomic_number(s)
    coords = np.array([xyz for _, xyz in records], dtype=np.float64) * scale
    return MoleculeSpec(symbols, coords, int(charge), int(spin))


def nuclear_repulsion_energy(spec: MoleculeSpec) -> float:
    z = spec.charges.astype(np.float64)
    r = spec.coords_bohr
    e = 0.0
    for i in range(len(z)):
        d = np.sqrt(((r[i + 1:] - r[i]) ** 2).sum(axis=1))
        e += float(np.sum(z[i] * z[i + 1:] / d))
    return e


# ---------------------------------------------------------------------------
# Deterministic workload geometries (all lengths in Angstrom before conversion)
# ---------------------------------------------------------------------------

WATER_OH = 0.96
WATER_HOH_DEG = 104.5


def _water_frame() -> np.ndarray:
    """O at the origin, the two H in the yz plane, bisector along +z."""
    half = math.radians(WATER_HOH_DEG) / 2.0
    return np.array([[0.0, 0.0, 0.0],
                     [0.0, WATER_OH * math.sin(half), WATER_OH * math.cos(half)],
                     [0.0, -WATER_OH * math.sin(half), WATER_OH * math.cos(half)]])


def _quaternion_rotation(q: np.ndarray) -> np.ndarray:
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def water_cluster(n: int, seed: int, box: float, min_oo: float = 2.6,
                  max_tries: int = 1_000_000) -> MoleculeSpec:
    """(H2O)_n: rigid monomers (O-H 0.96 A, HOH 104.5 deg) with uniformly random
    orientations (normalised Gaussian quaternions) and O positions drawn uniformly
    in [0, box]^3 A by rejection sampling with O-O >= min_oo A.  Atom order is
    O, H, H per monomer, monomers in placement order."""
    rng = np.random.default_rng(seed)
    frame = _water_frame()
    centres, atoms = [], []
    tries = 0
    while len(centres) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("water_cluster: rejection sampling did not converge")
        c = rng.uniform(0.0, box, size=3)
        if any(np.linalg.norm(c - o) < min_oo for o in centres):
            continue
        rot = _quaternion_rotation(rng.normal(size=4))
        centres.append(c)
        for sym, xyz in zip(("O", "H", "H"), frame @ rot.T + c):
            atoms.append((sym, xyz))
    return parse_molecule_spec(atoms, unit="Angstrom")


def water_monomer() -> MoleculeSpec:
    """Single water: O (0,0,0), H (0, +-0.757, 0.587) Angstrom, order O, H, H."""
    return parse_molecule_spec("O 0 0 0; H 0 0.757 0.587; H 0 -0.757 0.587")


def linear_alkane(n_carbon: int, cc: float = 1.54, ch: float = 1.09) -> MoleculeSpec:
    """All-trans C_nH_{2n+2} zig-zag with tetrahedral angles.  Carbons lie in the
    xy plane along x; each CH2 carries two H out of plane (+-z); the terminal
    carbons carry a third H continuing the zig-zag.  Atom order: all C, then H
    in carbon order."""
    tet = math.acos(-1.0 / 3.0)
    half = tet / 2.0
    dx = cc * math.sin(half)
    dy = cc * math.cos(half)
    carbons = [np.array([i * dx, (dy / 2.0) * (1 if i % 2 == 0 else -1), 0.0])
               for i in range(n_carbon)]
    hydrogens = []
    for i, c in enumerate(carbons):
        outward = np.array([0.0, 1.0 if i % 2 == 0 else -1.0, 0.0])
        # the two out-of-plane H of a tetrahedral CH2, bisector along `outward`
        for sz in (1.0, -1.0):
            v = outward * math.cos(half) + np.array([0.0, 0.0, sz * math.sin(half)])
            hydrogens.append(c + ch * v / np.linalg.norm(v))
        if i == 0 or i == n_carbon - 1:
            # the third terminal H sits where the next backbone carbon would be:
            # the C-C bond mirrored through the plane x = x_c
            nb = carbons[1] if i == 0 else carbons[-2]
            ref = np.array([-(nb - c)[0], (nb - c)[1], 0.0])
            ref = ref / np.linalg.norm(ref)
            hydrogens.append(c + ch * ref)
    atoms = [("C", c) for c in carbons] + [("H", h) for h in hydrogens]
    return parse_molecule_spec(atoms, unit="Angstrom")


def batch_water_clusters(count: int = 256, seed: int = 256, base_seed: int = 1000,
                         nmax: int = 8, box: float = 8.0) -> list:
    """Batch of (H2O)_n with n ~ U{1..nmax} from default_rng(seed); cluster i seeded base_seed+i."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, nmax + 1, size=count)
    return [water_cluster(int(n), base_seed + i, box) for i, n in enumerate(sizes)]


def to_xyz(spec: MoleculeSpec, comment: str = "") -> str:
    lines = [str(spec.natm), comment]
    for s, r in zip(spec.symbols, spec.coords_bohr / ANGSTROM_TO_BOHR):
        lines.append(f"{s} {r[0]:.12f} {r[1]:.12f} {r[2]:.12f}")
    return "\n".join(lines) + "\n"


def from_xyz(text: str) -> MoleculeSpec:
    lines = text.strip("\n").splitlines()
    n = int(lines[0])
    recs = []
    for ln in lines[2:2 + n]:
        p = ln.split()
        recs.append((p[0], tuple(float(v) for v in p[1:4])))
    return parse_molecule_spec(recs, unit="Angstrom")
