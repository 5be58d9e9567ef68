"""Density fitting (RI) of J and K on the MI355X engine.

Mirrors the reference's auxiliary-basis RI surface:
  * native two- and three-centre integrals in the compact layouts 2 / 3
    (GradSCFCompact -> density_fitting, csrc/density_fitting.cc:5-34, bound by
    backends/native/density_fitting.evaluate) -- here gscf_density_fitting on
    the GPU, never from four-centre ERIs;
  * AuxiliaryPlan / make_auxiliary_plan / AuxiliaryMetricFactor
    (src/gradscf/integrals/molecular/density_fitting.py:23-238): combined
    orbital + auxiliary tables, the metric factorised by Cholesky (eigen-
    decomposition with a lindep cut when it is not positive definite),
    whitened factors B = L^-1 (P|ab);
  * unpack_factors, build_jk_from_df, build_jk_from_df_orbitals,
    build_j_from_df (src/gradscf/integrals/molecular/jk.py:23-124): J and K
    from the factors by batched FP64 GEMMs (rocBLAS on the MI355X matrix
    cores), symmetrised like the reference.

Everything is a float64 torch tensor on the GPU; the factorisation runs on the
GPU as well (torch.linalg on rocSOLVER).
"""
from __future__ import annotations

import ctypes
import hashlib
from dataclasses import dataclass

import numpy as np
import torch

from . import _native
from .basis import BasisParameters, BasisTopology
from .jk import _device, _stream_handle
from .packing import pack


def _check_lindep(lindep):
    if not np.isfinite(lindep) or lindep <= 0:
        raise ValueError("lindep must be finite and positive.")


def evaluate(atm, bas, env, shape, *, layout: int, cart: bool, split: int) -> torch.Tensor:
    """Auxiliary integrals of combined tables (orbital shells [0, split), then
    auxiliary shells): layout 2 -> (P|Q) (m, m), layout 3 -> (ab|P) packed
    (m, n(n+1)/2); shape must be the layout's."""
    atm = np.ascontiguousarray(atm, dtype=np.int32)
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    out = torch.empty(tuple(int(s) for s in shape), dtype=torch.float64, device=_device())
    _native.check(_native.lib().gscf_density_fitting(
        atm.ctypes.data, len(atm), bas.ctypes.data, len(bas), env.ctypes.data, len(env), int(bool(cart)),
        int(layout), int(split), out.data_ptr(), out.numel(), _stream_handle()))
    return out


@dataclass(frozen=True)
class AuxiliaryMetricFactor:
    """Whitening of one auxiliary basis at one geometry: a lower Cholesky
    factor L of the metric (factors = L^-1 three) or, for a numerically
    dependent metric, the rows V_k / sqrt(w_k) of its eigenmodes above lindep
    (factors = matrix @ three)."""

    signature: str
    lindep: float
    method: str
    matrix: torch.Tensor

    def __post_init__(self):
        _check_lindep(self.lindep)
        m = self.matrix
        if (not isinstance(m, torch.Tensor) or m.dim() != 2 or not all(m.shape)
                or not bool(torch.isfinite(m).all())):
            raise ValueError("Auxiliary metric factor must be a finite nonempty matrix.")
        if self.method == "cholesky":
            if (m.shape[0] != m.shape[1] or bool((torch.diagonal(m) <= 0).any())
                    or bool((torch.triu(m, 1) != 0).any())):
                raise ValueError("Auxiliary Cholesky factor must be lower triangular.")
        elif self.method != "eigh" or m.shape[0] > m.shape[1]:
            raise ValueError("Unknown auxiliary metric factorization or rank.")

    def whiten(self, three: torch.Tensor) -> torch.Tensor:
        if three.dim() != 2 or three.shape[0] != self.matrix.shape[1] or not bool(torch.isfinite(three).all()):
            raise ValueError("Three-center integrals must be finite and match the auxiliary metric.")
        if self.method == "cholesky":
            factors = torch.linalg.solve_triangular(self.matrix, three, upper=False)
        else:
            factors = self.matrix @ three
        if not bool(torch.isfinite(factors).all()):
            raise ValueError("Whitened auxiliary factors must be finite.")
        return factors


def _factor_metric(metric: torch.Tensor, signature: str, lindep: float) -> AuxiliaryMetricFactor:
    if not bool(torch.isfinite(metric).all()):
        raise ValueError("Auxiliary Coulomb metric must be finite.")
    chol, info = torch.linalg.cholesky_ex(metric)
    if int(info) == 0:
        return AuxiliaryMetricFactor(signature, lindep, "cholesky", chol)
    w, v = torch.linalg.eigh(metric)
    keep = w > lindep
    if not bool(keep.any()):
        raise ValueError("Auxiliary Coulomb metric has no positive independent modes.")
    return AuxiliaryMetricFactor(signature, lindep, "eigh", (v[:, keep] / torch.sqrt(w[keep])).T.contiguous())


@dataclass(frozen=True)
class AuxiliaryPlan:
    orbital: BasisTopology
    auxiliary: BasisTopology
    combined: BasisTopology

    @property
    def split(self) -> int:
        return len(self.orbital.angular_momenta)

    def _pack(self, parameters: BasisParameters, auxiliary_parameters: BasisParameters):
        p, a = parameters, auxiliary_parameters
        combined = BasisParameters(tuple(p.exponents) + tuple(a.exponents),
                                   tuple(p.coefficients) + tuple(a.coefficients),
                                   np.concatenate((np.asarray(p.centers), np.asarray(a.centers))),
                                   np.asarray(p.nuclear_coords))
        return pack(self.combined, combined)

    def _signature(self, parameters, auxiliary_parameters) -> str:
        """Auxiliary basis and geometry (not the orbital exponents/coefficients)."""
        a = auxiliary_parameters
        digest = hashlib.sha256(repr(self.auxiliary).encode())
        digest.update(repr(self.orbital.nuclear_charges).encode())

        def add(value):
            arr = np.ascontiguousarray(value, dtype=np.float64)
            if not np.all(np.isfinite(arr)):
                raise ValueError("Auxiliary basis and geometry parameters must be finite.")
            digest.update(repr(arr.shape).encode())
            digest.update(arr.tobytes())
            return arr

        for alpha in a.exponents:
            if np.any(add(alpha) <= 0):
                raise ValueError("Auxiliary exponents must be positive.")
        for c in a.coefficients:
            add(c)
        for coords in (a.centers, a.nuclear_coords, parameters.nuclear_coords):
            add(coords)
        centers = np.asarray(parameters.centers)
        if centers.ndim != 2 or centers.shape[1] != 3:
            raise ValueError("Orbital geometry centers must have shape (nshell,3).")
        add(np.unique(centers, axis=0))
        return digest.hexdigest()

    def evaluate(self, parameters, auxiliary_parameters):
        """(metric (m, m), three-centre (m, n(n+1)/2)) on the GPU."""
        atm, bas, env = self._pack(parameters, auxiliary_parameters)
        n, m = self.orbital.nao(), self.auxiliary.nao()
        kw = dict(cart=self.orbital.cart, split=self.split)
        metric = evaluate(atm, bas, env, (m, m), layout=2, **kw)
        three = evaluate(atm, bas, env, (m, n * (n + 1) // 2), layout=3, **kw)
        return metric, three

    def metric_factor(self, parameters, auxiliary_parameters, *, lindep=1e-12) -> AuxiliaryMetricFactor:
        """Factorise the two-centre metric once, for repeated three-centre
        evaluations at the same auxiliary basis and geometry."""
        _check_lindep(lindep)
        signature = self._signature(parameters, auxiliary_parameters)
        atm, bas, env = self._pack(parameters, auxiliary_parameters)
        m = self.auxiliary.nao()
        metric = evaluate(atm, bas, env, (m, m), layout=2, cart=self.orbital.cart, split=self.split)
        return _factor_metric(metric, signature, lindep)

    def factors(self, parameters, auxiliary_parameters, *, lindep=1e-12, metric_factor=None) -> torch.Tensor:
        """Whitened packed factors (rank, n(n+1)/2); with metric_factor only
        the three-centre integrals are evaluated."""
        _check_lindep(lindep)
        signature = self._signature(parameters, auxiliary_parameters)
        if metric_factor is None:
            metric, three = self.evaluate(parameters, auxiliary_parameters)
            metric_factor = _factor_metric(metric, signature, lindep)
            return metric_factor.whiten(three)
        if (not isinstance(metric_factor, AuxiliaryMetricFactor) or metric_factor.signature != signature
                or metric_factor.lindep != lindep or metric_factor.matrix.shape[1] != self.auxiliary.nao()):
            raise ValueError("Cached auxiliary metric signature, geometry or lindep differs.")
        atm, bas, env = self._pack(parameters, auxiliary_parameters)
        n, m = self.orbital.nao(), self.auxiliary.nao()
        three = evaluate(atm, bas, env, (m, n * (n + 1) // 2), layout=3, cart=self.orbital.cart, split=self.split)
        return metric_factor.whiten(three)


def make_auxiliary_plan(orbital: BasisTopology, auxiliary: BasisTopology) -> AuxiliaryPlan:
    if orbital.cart != auxiliary.cart:
        raise ValueError("Orbital and auxiliary angular representations must match.")
    top = BasisTopology(orbital.angular_momenta + auxiliary.angular_momenta,
                        orbital.primitive_counts + auxiliary.primitive_counts,
                        orbital.contraction_counts + auxiliary.contraction_counts,
                        orbital.nuclear_charges, orbital.cart)
    return AuxiliaryPlan(orbital, auxiliary, top)


def unpack_factors(packed: torch.Tensor, nao: int) -> torch.Tensor:
    """(rank, n(n+1)/2) -> symmetric (rank, n, n); never a four-index tensor."""
    if packed.dim() != 2 or packed.shape[1] != nao * (nao + 1) // 2:
        raise ValueError("Packed RI factor shape mismatch.")
    i, j = np.tril_indices(nao)
    ii = torch.as_tensor(i, device=packed.device)
    jj = torch.as_tensor(j, device=packed.device)
    out = torch.zeros((packed.shape[0], nao, nao), dtype=packed.dtype, device=packed.device)
    out[:, ii, jj] = packed
    out[:, jj, ii] = packed
    return out


def build_jk_from_df(df_factors: torch.Tensor, density: torch.Tensor):
    """J and K of density (n, n) from unpacked factors B (Q, n, n):
    J = sum_Q (B_Q . D) B_Q, K = sum_Q B_Q D B_Q^T, each symmetrised."""
    if df_factors.shape[0] == 0:
        z = torch.zeros_like(density)
        return z, z
    nq, n = df_factors.shape[0], df_factors.shape[1]
    flat = df_factors.reshape(nq, n * n)
    rho = flat @ density.reshape(n * n)                               # (Q|D)
    j = (rho @ flat).reshape(n, n)
    # K = sum_Q (B_Q D) B_Q^T as one (n, Qn) x (Qn, n) GEMM
    projected = torch.matmul(df_factors, density)                    # B_Q D
    x = projected.transpose(0, 1).reshape(n, nq * n)
    y = df_factors.transpose(0, 1).reshape(n, nq * n)
    k = x @ y.T
    return 0.5 * (j + j.T), 0.5 * (k + k.T)


def build_jk_from_df_orbitals(df_factors, density, mo_coeff, mo_occ, *, nocc: int):
    """J from the density, K from occupied orbitals: K = sum_Q (B_Q C~)(B_Q C~)^T
    with C~ = C_occ sqrt(occ), an (rank, n, nocc) intermediate instead of
    (rank, n, n)."""
    if df_factors.shape[0] == 0:
        z = torch.zeros_like(density)
        return z, z
    nocc = int(nocc)
    if nocc <= 0 or nocc > mo_coeff.shape[-1]:
        raise ValueError("nocc must be in the range [1, nmo].")
    nq, n = df_factors.shape[0], df_factors.shape[1]
    flat = df_factors.reshape(nq, n * n)
    j = ((flat @ density.reshape(n * n)) @ flat).reshape(n, n)
    c = mo_coeff[:, :nocc] * torch.sqrt(torch.clamp(mo_occ[:nocc], min=0.0))[None, :]
    b = torch.matmul(df_factors, c)                                 # (Q, n, nocc)
    x = b.transpose(0, 1).reshape(n, nq * nocc)
    k = x @ x.T
    return 0.5 * (j + j.T), 0.5 * (k + k.T)


def build_j_from_df(df_factors: torch.Tensor, density: torch.Tensor) -> torch.Tensor:
    if df_factors.shape[0] == 0:
        return torch.zeros_like(density)
    nq, n = df_factors.shape[0], df_factors.shape[1]
    flat = df_factors.reshape(nq, n * n)
    j = ((flat @ density.reshape(n * n)) @ flat).reshape(n, n)
    return 0.5 * (j + j.T)


def compact(atm, bas, env, shape, *, layout: int, cart: bool, split: int = 0) -> torch.Tensor:
    """The reference's compact target (csrc/compact.cc:5-21) in one call:
    layouts 0 / 1 (packed four-centre ERIs) and 2 / 3 (auxiliary integrals)."""
    atm = np.ascontiguousarray(atm, dtype=np.int32)
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    out = torch.empty(tuple(int(s) for s in shape), dtype=torch.float64, device=_device())
    _native.check(_native.lib().gscf_compact(
        atm.ctypes.data, len(atm), bas.ctypes.data, len(bas), env.ctypes.data, len(env), int(layout),
        int(bool(cart)), int(split), out.data_ptr(), out.numel(), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return out
