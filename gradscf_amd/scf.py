"""Restricted Hartree-Fock SCF on the MI355X engine (direct J/K call site).

Follows the reference SCF semantics:
  * Loewdin X = S^{-1/2}, eigenvalues clipped at 1e-10
    (src/gradscf/solvers/eigen/spectral.py:38-42);
  * core-Hamiltonian guess, D = C occ C^T with occ = 2 (src/gradscf/scf/rks.py:1121-1152);
  * each cycle: DIIS from cycle >= 1 on the error C0^T (F D S - S D F) C0 with C0
    the core-guess orbitals, 8-deep ring, extrapolation once two entries exist
    (src/gradscf/scf/rks.py:166-170, 896-945; src/gradscf/solvers/nonlinear/diis.py:18-103);
    diagonalise; J/K on the density change D_new - D_old added to the previous
    J/K (the incremental build of src/gradscf/scf/rks.py:309-322);
    F = H + J - K/2, E = tr(DH) + tr(DJ)/2 - tr(DK)/4 + E_nn (src/gradscf/scf/energy.py:23-31);
  * convergence |dE| < conv_tol, rms(dD) < 1e-8, ||2 Cv^T F Co|| < 1e-7 after
    cycle 0 (src/gradscf/scf/rks.py:787-846, src/gradscf/scf/convergence.py:7-24).
All linear algebra stays FP64 on the GPU; J/K come from the HIP engine.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .basis import prepare_basis
from .jk import DirectJKPlan
from .linalg import eigh
from .molecule import MoleculeSpec, parse_molecule_spec
from .packing import pack

DIIS_SPACE = 8


@dataclass
class RHFResult:
    energy: float
    converged: bool
    cycles: int
    density: torch.Tensor
    mo_coeff: torch.Tensor
    mo_energy: torch.Tensor
    fock: torch.Tensor
    j: torch.Tensor
    k: torch.Tensor
    nuclear_repulsion: float
    timings: dict = field(default_factory=dict)


def orthogonalizer(s: torch.Tensor, eps: float = 1e-10) -> torch.Tensor:
    w, v = eigh(s)
    return (v * torch.clamp(w, min=eps).rsqrt()) @ v.T


def _diagonalize(f, x):
    fo = x.T @ f @ x
    fo = 0.5 * (fo + fo.T)
    e, c = eigh(fo)
    return e, x @ c


def _diis_solve(fock_hist, gram, count):
    """Extrapolated Fock from the ring history; `gram` = E E^T of the error ring,
    maintained one row per push (a (8 x n^2)(n^2 x 8) product is a poor fit for
    a GEMM: tiny output, huge reduction)."""
    m = fock_hist.shape[0]
    dev, dt = fock_hist.device, fock_hist.dtype
    valid = (torch.arange(m, device=dev) < count).to(dt)
    scale = torch.max(torch.abs(gram))
    scale = torch.where(scale > 0, scale, torch.ones_like(scale))
    gram = gram / scale
    top = gram * (valid[:, None] * valid[None, :]) + torch.diag(valid * 50 * torch.finfo(dt).eps + (1 - valid))
    b = torch.zeros((m + 1, m + 1), dtype=dt, device=dev)
    b[:m, :m] = top
    b[:m, m] = -valid
    b[m, :m] = -valid
    rhs = torch.zeros(m + 1, dtype=dt, device=dev)
    rhs[m] = -1.0
    coeff = torch.linalg.solve(b, rhs)[:m] * valid
    return torch.tensordot(coeff, fock_hist, dims=([0], [0]))


def rhf(plan: DirectJKPlan, nelectron: int, enuc: float, *, conv_tol=1e-10, conv_tol_density=1e-8,
        conv_tol_grad=1e-7, max_cycle=80, jk_flags=None, incore=False) -> RHFResult:
    """Run RHF with the direct J/K engine (`plan` already holds the basis and screening).

    incore: evaluate the plan's integrals once and digest stored blocks in
    every cycle (DirectJKPlan.incore; released again at the end unless the
    plan already held them)."""
    own_incore = bool(incore) and not getattr(plan, "incore_enabled", False)
    if own_incore:
        plan.incore(True)
    try:
        return _rhf(plan, nelectron, enuc, conv_tol, conv_tol_density, conv_tol_grad, max_cycle, jk_flags)
    finally:
        if own_incore:
            plan.incore(False)


def _rhf(plan, nelectron, enuc, conv_tol, conv_tol_density, conv_tol_grad, max_cycle, jk_flags) -> RHFResult:
    jk_flags = dict(jk_flags or {})
    t0 = time.perf_counter()
    s, t, v = plan.int1e()
    h = t + v
    n = h.shape[0]
    nocc = nelectron // 2
    if nelectron % 2 or nocc <= 0 or nocc > n:
        raise ValueError("RHF requires an even number of electrons and enough orbitals.")
    occ = torch.zeros(n, dtype=torch.float64, device=h.device)
    occ[:nocc] = 2.0
    torch.cuda.synchronize()
    t_int = time.perf_counter()

    def jk(d):
        out = plan.jk(d[None], symmetric=True, validate=False, **jk_flags)
        return out[0, 0], out[1, 0]

    x = orthogonalizer(s)
    mo_e, c = _diagonalize(h, x)
    c0 = c
    d = (c * occ) @ c.T
    j, k = jk(d)
    fock = h + j - 0.5 * k
    energy = 0.0
    fh = torch.zeros((DIIS_SPACE, n, n), dtype=torch.float64, device=h.device)
    eh = torch.zeros((DIIS_SPACE, n * n), dtype=torch.float64, device=h.device)
    gram = torch.zeros((DIIS_SPACE, DIIS_SPACE), dtype=torch.float64, device=h.device)
    head = count = 0
    converged, cycles = False, 0
    n_jk = 1
    for cycle in range(max_cycle):
        fock_eff = fock
        if cycle >= 1:
            err = c0.T @ (fock @ d @ s - s @ d @ fock) @ c0
            fh[head] = fock
            eh[head] = err.reshape(-1)
            row = (eh * eh[head]).sum(dim=1)
            gram[head, :] = row
            gram[:, head] = row
            head = (head + 1) % DIIS_SPACE
            count = min(count + 1, DIIS_SPACE)
            if count >= 2:
                fock_eff = _diis_solve(fh, gram, count)
        mo_e, c = _diagonalize(fock_eff, x)
        d_new = (c * occ) @ c.T
        dj, dk = jk(d_new - d)
        n_jk += 1
        j, k = j + dj, k + dk
        fock_new = h + j - 0.5 * k
        # energy, density change and orbital gradient in one host transfer
        e_el, rms, grad = torch.stack([
            torch.sum(d_new * h) + 0.5 * torch.sum(d_new * j) - 0.25 * torch.sum(d_new * k),
            torch.sqrt(torch.mean((d_new - d) ** 2)),
            torch.linalg.norm(2.0 * c[:, nocc:].T @ fock_new @ c[:, :nocc])]).tolist()
        e_new = e_el + enuc
        done = abs(e_new - energy) < conv_tol and rms < conv_tol_density and grad < conv_tol_grad and cycle > 0
        d, fock, energy = d_new, fock_new, e_new
        cycles = cycle + 1
        if done:
            converged = True
            break
    torch.cuda.synchronize()
    t_end = time.perf_counter()
    return RHFResult(energy, converged, cycles, d, c, mo_e, fock, j, k, enuc,
                     {"integrals_s": t_int - t0, "scf_s": t_end - t_int, "jk_builds": n_jk})


class RHF:
    """Facade mirroring the reference's RHF(mol, jk_backend='direct').run()
    (src/gradscf/scf/facade.py:348-478) for the direct backend."""

    def __init__(self, atom, basis, *, unit="Angstrom", charge=0, cart=True, conv_tol=1e-10,
                 direct_scf_tol=0.0, max_cycle=80, mixed_threshold=None):
        self.spec = atom if isinstance(atom, MoleculeSpec) else parse_molecule_spec(atom, unit=unit, charge=charge)
        self.basis = basis
        self.cart = cart
        self.conv_tol = conv_tol
        self.direct_scf_tol = direct_scf_tol
        self.max_cycle = max_cycle
        self.mixed_threshold = mixed_threshold
        topo, par = prepare_basis(self.spec, basis, cart=cart)
        self.atm, self.bas, self.env = pack(topo, par)
        self.plan = DirectJKPlan(self.atm, self.bas, self.env, cart=cart, cutoff=direct_scf_tol)
        self.result: RHFResult | None = None

    def run(self) -> RHFResult:
        flags = {}
        if self.mixed_threshold is not None:
            flags["mixed_threshold"] = self.mixed_threshold
        self.result = rhf(self.plan, self.spec.nelectron, self.spec.nuclear_repulsion(), conv_tol=self.conv_tol,
                          max_cycle=self.max_cycle, jk_flags=flags)
        return self.result

    def kernel(self) -> float:
        return self.run().energy
