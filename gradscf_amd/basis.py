"""Named basis sets, basis topology/parameters and libcint-convention normalization.

Reference semantics followed:
  * named-basis parsing order (per atom: all s shells, then p, then d, in file
    order; SP blocks split into an s and a p shell):
    src/gradscf/integrals/basis/data/__init__.py:176-205
  * prepare_basis (raw, unnormalised per-shell exponents/coefficients, one
    centre per shell): src/gradscf/integrals/basis/library.py:5-51
  * radial primitive norm and per-column renormalisation of contracted shells:
    src/gradscf/integrals/basis/normalization.py:50-63

The basis numbers (exponents and contraction coefficients) are the published
STO-3G, 6-31G, 6-31G* and cc-pVDZ sets and the cc-pVDZ-JKFIT and
def2-universal-JKFIT auxiliary sets for H, He, C, N, O, stored as JSON under
basis_data/ in the raw [l, [exponent, c_1, ..., c_nctr], ...] layout
(basis_data/extract.py regenerates them from the reference's snapshot).
"""
from __future__ import annotations

from dataclasses import dataclass
import json
import math
from functools import lru_cache
from pathlib import Path

import numpy as np

from .molecule import MoleculeSpec, parse_molecule_spec, std_symbol

_DATA = Path(__file__).with_name("basis_data")

_ALIASES = {
    "sto3g": "sto-3g",
    "631g": "6-31g",
    "631g*": "6-31gs",
    "631gs": "6-31gs",
    "ccpvdz": "cc-pvdz",
    "ccpvdzjkfit": "cc-pvdz-jkfit",
    "def2universaljkfit": "def2-universal-jkfit",
}


def _format_name(name: str) -> str:
    return str(name).lower().replace("-", "").replace("_", "").replace(" ", "")


@lru_cache(maxsize=None)
def _load_file(stem: str) -> dict:
    return json.loads((_DATA / f"{stem}.json").read_text())


def load_basis(name: str, symbol: str) -> list:
    """Raw shells [[l, [exp, c...], ...], ...] of a named basis for one element."""
    key = _format_name(name)
    if key not in _ALIASES:
        raise KeyError(f"Basis {name!r} is not available (have: sto-3g, 6-31g, 6-31g*, cc-pvdz, "
                       "cc-pvdz-jkfit, def2-universal-jkfit).")
    data = _load_file(_ALIASES[key])
    sym = std_symbol(symbol)
    if sym not in data:
        raise KeyError(f"Basis {name!r} has no data for element {sym!r}.")
    return json.loads(json.dumps(data[sym]))


@dataclass(frozen=True)
class BasisTopology:
    angular_momenta: tuple
    primitive_counts: tuple
    contraction_counts: tuple
    nuclear_charges: tuple
    cart: bool

    @property
    def nshell(self) -> int:
        return len(self.angular_momenta)

    def nao(self) -> int:
        n = 0
        for l, nc in zip(self.angular_momenta, self.contraction_counts):
            n += nc * ((l + 1) * (l + 2) // 2 if self.cart else 2 * l + 1)
        return n


@dataclass(frozen=True)
class BasisParameters:
    exponents: tuple          # per shell: (nprim,)
    coefficients: tuple       # per shell: (nprim, nctr), raw (unnormalised)
    centers: np.ndarray       # (nshell, 3) Bohr
    nuclear_coords: np.ndarray  # (natm, 3) Bohr


def prepare_basis(atom, basis, *, unit="Angstrom", charge=0, spin=0, cart=True):
    """Parse a named or explicit basis into (topology, parameters)."""
    spec = atom if isinstance(atom, MoleculeSpec) else parse_molecule_spec(
        atom, unit=unit, charge=charge, spin=spin)
    ls, nprim, nctr, exps, coefs, centers = [], [], [], [], [], []
    for i, symbol in enumerate(spec.symbols):
        value = basis[symbol] if isinstance(basis, dict) else basis
        blocks = load_basis(value, symbol) if isinstance(value, str) else value
        for block in blocks:
            l, rows = int(block[0]), block[1:]
            if not rows or not isinstance(rows[0], (list, tuple)):
                raise ValueError("Expected [l, [exponent, coefficients...], ...] shells.")
            data = np.asarray(rows, dtype=np.float64)
            if data.ndim != 2 or data.shape[1] < 2:
                raise ValueError("Each primitive requires an exponent and coefficients.")
            ls.append(l)
            nprim.append(data.shape[0])
            nctr.append(data.shape[1] - 1)
            exps.append(data[:, 0].copy())
            coefs.append(data[:, 1:].copy())
            centers.append(spec.coords_bohr[i].copy())
    topo = BasisTopology(tuple(ls), tuple(nprim), tuple(nctr),
                         tuple(int(z) for z in spec.charges), bool(cart))
    params = BasisParameters(tuple(exps), tuple(coefs), np.array(centers).reshape(-1, 3),
                             np.asarray(spec.coords_bohr, dtype=np.float64))
    return topo, params


def radial_primitive_norm(l: int, exponents) -> np.ndarray:
    """1/sqrt(int_0^inf r^{2l+2} exp(-2 a r^2) dr)."""
    a = np.asarray(exponents, dtype=np.float64)
    n = l + 1.5
    return 1.0 / np.sqrt(math.gamma(n) / (2.0 * (2.0 * a) ** n))


def normalized_shell_coefficients(l: int, exponents, coefficients) -> np.ndarray:
    """Raw (nprim, nctr) coefficients -> normalised libcint environment values.

    Each column is scaled so that the contracted radial function r^l sum_p c_p
    exp(-a_p r^2) has unit norm (no s/p angular factor: libcint applies those
    inside its integral kernels)."""
    a = np.asarray(exponents, dtype=np.float64)
    c = np.asarray(coefficients, dtype=np.float64)
    if c.ndim == 1:
        c = c[:, None]
    w = c * radial_primitive_norm(l, a)[:, None]
    metric = math.gamma(l + 1.5) / (2.0 * (a[:, None] + a[None, :]) ** (l + 1.5))
    norm2 = np.einsum("pi,pq,qi->i", w, metric, w)
    return w / np.sqrt(norm2)[None, :]


def normalized_shell_coefficients_torch(l: int, exponents, coefficients):
    """normalized_shell_coefficients in torch operations, so that gradients
    with respect to the normalised values chain to raw exponents and
    coefficients (the reference differentiates the same map,
    src/gradscf/integrals/basis/normalization.py:57-63)."""
    import torch
    a = torch.as_tensor(exponents, dtype=torch.float64)
    c = torch.as_tensor(coefficients, dtype=torch.float64, device=a.device)
    if c.dim() == 1:
        c = c[:, None]
    n = l + 1.5
    w = c * (1.0 / torch.sqrt(math.gamma(n) / (2.0 * (2.0 * a) ** n)))[:, None]
    metric = math.gamma(n) / (2.0 * (a[:, None] + a[None, :]) ** n)
    norm2 = torch.einsum("pi,pq,qi->i", w, metric, w)
    return w / torch.sqrt(norm2)[None, :]


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def nsph(l: int) -> int:
    return 2 * l + 1
