"""Batched RHF over many independent molecules of uneven size.

Each molecule follows exactly the single-molecule driver in gradscf_amd.scf
(the reference semantics cited there: Loewdin orthogonaliser, core guess,
8-deep DIIS on C0^T (FDS - SDF) C0, incremental J/K on D_new - D_old,
convergence on |dE|, rms(dD) and the orbital gradient).  What changes is the
schedule:

  * one packed J/K call per SCF iteration covers every molecule (a batched
    plan: one launch per quartet class for the whole batch), optionally
    replayed from a captured HIP graph;
  * every molecule's matrices sit in one (nmol, N, N) batch padded to the
    largest AO count N (identity overlap in the padding, masks for occupied
    and virtual orbitals), so DIIS extrapolation and convergence tests run as
    single batched FP64 tensor operations over all molecules;
  * the diagonalisations solve each molecule's own block, molecules of one
    AO count together, the counts concurrently on several streams
    (linalg.BatchedEigh: rocSOLVER's batched solver, no host
    synchronisation), and skip molecules that have converged;
  * a molecule that has converged is frozen (its density change is zero from
    then on), so its energy is the one an independent run would stop at.

The reference has no multi-molecule SCF; its batch rule maps molecules one
call at a time (src/gradscf/integrals/backends/native/jk.py:253-317).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .jk import DirectJKPlan
from .linalg import BatchedEigh

DIIS_SPACE = 8


@dataclass
class BatchRHFResult:
    energies: np.ndarray          # (nmol,) Eh
    converged: np.ndarray         # (nmol,) bool
    cycles: np.ndarray            # (nmol,) SCF cycles until convergence (or max_cycle)
    density: torch.Tensor         # packed (nn_total,) final densities
    iteration_s: list             # wall time of every SCF iteration (all molecules)
    timings: dict = field(default_factory=dict)


class _Padded:
    """Every molecule's n_m x n_m matrices in one (nmol, N, N) batch, N the
    largest AO count, zeros outside each molecule's block; gather / scatter
    between that batch and the packed layout of a batched plan."""

    def __init__(self, mol_n, offsets, device):
        n = np.asarray(mol_n, dtype=np.int64)
        self.k, self.N = len(n), int(n.max())
        src, dst = [], []
        for m, (nm, off) in enumerate(zip(n, np.asarray(offsets, dtype=np.int64))):
            ii, jj = np.meshgrid(np.arange(nm), np.arange(nm), indexing="ij")
            src.append(off + (ii * nm + jj).ravel())
            dst.append(m * self.N * self.N + (ii * self.N + jj).ravel())
        self.src = torch.as_tensor(np.concatenate(src), device=device)
        self.dst = torch.as_tensor(np.concatenate(dst), device=device)
        ar = torch.arange(self.N, device=device)
        nt = torch.as_tensor(n, device=device)
        self.valid = (ar[None, :] < nt[:, None]).to(torch.float64)      # (k, N) AO rows of each molecule
        self.n = nt.to(torch.float64)

    def gather(self, packed):
        out = torch.zeros(self.k * self.N * self.N, dtype=packed.dtype, device=packed.device)
        out[self.dst] = packed[self.src]
        return out.view(self.k, self.N, self.N)

    def scatter(self, packed, mats):
        packed[self.src] = mats.reshape(-1)[self.dst]


def _orthogonalizer(s, eig, eps=1e-10):
    """Loewdin S^{-1/2} of every molecule's block (the identity in the
    padding: unit eigenvalues there)."""
    w, v = eig(s, e_out=torch.ones((eig.k, eig.N), dtype=s.dtype, device=s.device))
    return (v * torch.clamp(w, min=eps).rsqrt()[:, None, :]) @ v.transpose(1, 2)


def _diagonalize(f, x, eig, c_buf=None, active_h=None):
    """Eigenvectors of x^T f x per molecule block (each molecule's in its
    leading columns, the identity in the padding); molecules outside active_h
    keep the columns c_buf held."""
    fo = x.transpose(1, 2) @ f @ x
    fo = 0.5 * (fo + fo.transpose(1, 2))
    _, c = eig(fo, c_out=c_buf, active_h=active_h)
    return c, x @ c


def _diis_solve(fock_hist, gram, count):
    """Batched form of the single-molecule DIIS solve: fock_hist (k, m, n, n),
    gram (k, m, m) = E E^T of each molecule's error ring."""
    k, m = gram.shape[0], gram.shape[1]
    dev, dt = gram.device, gram.dtype
    valid = (torch.arange(m, device=dev) < count).to(dt)
    scale = torch.amax(torch.abs(gram), dim=(1, 2))
    scale = torch.where(scale > 0, scale, torch.ones_like(scale))
    g = gram / scale[:, None, None]
    top = g * (valid[:, None] * valid[None, :]) + torch.diag(valid * 50 * torch.finfo(dt).eps + (1 - valid))
    b = torch.zeros((k, m + 1, m + 1), dtype=dt, device=dev)
    b[:, :m, :m] = top
    b[:, :m, m] = -valid
    b[:, m, :m] = -valid
    rhs = torch.zeros((k, m + 1), dtype=dt, device=dev)
    rhs[:, m] = -1.0
    coeff = torch.linalg.solve_ex(b, rhs)[0][:, :m] * valid
    k_, m_, n_ = fock_hist.shape[0], fock_hist.shape[1], fock_hist.shape[2]
    return torch.bmm(coeff[:, None, :], fock_hist.reshape(k_, m_, n_ * n_)).reshape(k_, n_, n_)


class _GraphJK:
    """J/K of a packed density replayed from a captured HIP graph: the build
    issues kernel launches only (no host sync, memset or allocation once the
    workspace exists), so the whole per-iteration J/K is one graph launch."""

    def __init__(self, plan: DirectJKPlan, mixed_threshold):
        self.plan = plan
        self.thr = mixed_threshold
        dev = torch.device("cuda", torch.cuda.current_device())
        self.d = torch.zeros((1, plan.nn_total), dtype=torch.float64, device=dev)
        self.out = torch.empty((2, 1, plan.nn_total), dtype=torch.float64, device=dev)
        self._run()  # allocates the engine workspace outside the capture
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._run()
        torch.cuda.synchronize()

    def _run(self):
        self.plan.jk(self.d, symmetric=True, validate=False, out=self.out, mixed_threshold=self.thr)

    def __call__(self, packed):
        self.d[0].copy_(packed)
        self.graph.replay()
        return self.out[0, 0].clone(), self.out[1, 0].clone()


def rhf_batch(plan: DirectJKPlan, nelectrons, enucs, *, conv_tol=1e-10, conv_tol_density=1e-8,
              conv_tol_grad=1e-7, max_cycle=80, mixed_threshold=None, use_graph=True,
              incore=False) -> BatchRHFResult:
    """RHF for every molecule of a batched plan (packed layout); nelectrons and
    enucs give each molecule's electron count and nuclear repulsion.

    incore: evaluate the plan's integrals once and digest stored blocks in
    every iteration (DirectJKPlan.incore; released again at the end unless the
    plan already held them)."""
    own_incore = bool(incore) and not getattr(plan, "incore_enabled", False)
    if own_incore:
        plan.incore(True)
    try:
        return _rhf_batch(plan, nelectrons, enucs, conv_tol, conv_tol_density, conv_tol_grad, max_cycle,
                          mixed_threshold, use_graph)
    finally:
        if own_incore:
            plan.incore(False)


def _rhf_batch(plan, nelectrons, enucs, conv_tol, conv_tol_density, conv_tol_grad, max_cycle, mixed_threshold,
               use_graph) -> BatchRHFResult:
    nmol = plan.nmol
    nelectrons = np.asarray(nelectrons, dtype=np.int64)
    enucs = torch.as_tensor(np.asarray(enucs, dtype=np.float64))
    if nelectrons.shape != (nmol,) or enucs.shape != (nmol,):
        raise ValueError("one electron count and one nuclear repulsion per molecule")
    if np.any(nelectrons % 2) or np.any(nelectrons <= 0) or np.any(nelectrons // 2 > plan.mol_n):
        raise ValueError("RHF requires an even number of electrons and enough orbitals.")
    t0 = time.perf_counter()
    s, t, v = plan.int1e()
    h = t + v
    dev = h.device
    enuc = enucs.to(dev)
    B = _Padded(plan.mol_n, plan.mol_offset, dev)
    k, N = B.k, B.N
    ar = torch.arange(N, device=dev)
    nocc = torch.as_tensor(nelectrons // 2, device=dev)
    occ = (ar[None, :] < nocc[:, None]).to(torch.float64)          # (k, N) occupied MO columns
    vir = B.valid - occ                                             # virtual MO columns
    pad = 1.0 - B.valid
    S = B.gather(s) + torch.diag_embed(pad)                         # identity in the padding
    H = B.gather(h)
    torch.cuda.synchronize()
    t_int = time.perf_counter()

    if use_graph:
        jk_graph = _GraphJK(plan, mixed_threshold)

        def jk(packed):
            return jk_graph(packed)
    else:
        def jk(packed):
            out = plan.jk(packed[None], symmetric=True, validate=False, mixed_threshold=mixed_threshold)
            return out[0, 0], out[1, 0]

    eig = BatchedEigh(plan.mol_n, dev)
    X = _orthogonalizer(S, eig)
    c_buf, C0 = _diagonalize(H, X, eig)
    Co = C0 * occ[:, None, :]
    D = 2.0 * Co @ Co.transpose(1, 2)
    d_packed = torch.zeros(plan.nn_total, dtype=torch.float64, device=dev)
    B.scatter(d_packed, D)
    fh = torch.zeros((k, DIIS_SPACE, N, N), dtype=torch.float64, device=dev)
    eh = torch.zeros((k, DIIS_SPACE, N * N), dtype=torch.float64, device=dev)
    gram = torch.zeros((k, DIIS_SPACE, DIIS_SPACE), dtype=torch.float64, device=dev)
    energy = torch.zeros(k, dtype=torch.float64, device=dev)
    active = torch.ones(k, dtype=torch.bool, device=dev)
    cycles = np.zeros(k, dtype=np.int64)
    j_packed, k_packed = jk(d_packed)
    J, K = B.gather(j_packed), B.gather(k_packed)
    F = H + J - 0.5 * K
    head = count = 0
    active_h = None
    iteration_s = []
    n_jk = 1
    torch.cuda.synchronize()
    for cycle in range(max_cycle):
        ti = time.perf_counter()
        f_eff = F
        if cycle >= 1:
            err = C0.transpose(1, 2) @ (F @ D @ S - S @ D @ F) @ C0
            fh[:, head] = F
            eh[:, head] = err.reshape(k, -1)
            row = torch.bmm(eh, eh[:, head, :, None])[:, :, 0]
            gram[:, head, :] = row
            gram[:, :, head] = row
        if cycle >= 1 and min(count + 1, DIIS_SPACE) >= 2:
            f_eff = _diis_solve(fh, gram, min(count + 1, DIIS_SPACE))
        c_buf, C = _diagonalize(f_eff, X, eig, c_buf, active_h)
        Co = C * occ[:, None, :]
        D_new = 2.0 * Co @ Co.transpose(1, 2)
        D_new = torch.where(active[:, None, None], D_new, D)
        dd_packed = torch.zeros_like(d_packed)
        B.scatter(dd_packed, D_new - D)
        if cycle >= 1:
            head = (head + 1) % DIIS_SPACE
            count = min(count + 1, DIIS_SPACE)
        dj_packed, dk_packed = jk(dd_packed)
        n_jk += 1
        J = J + B.gather(dj_packed)
        K = K + B.gather(dk_packed)
        F_new = H + J - 0.5 * K
        e_new = (D_new * (H + 0.5 * J - 0.25 * K)).sum(dim=(1, 2)) + enuc
        rms = torch.sqrt(((D_new - D) ** 2).sum(dim=(1, 2)) / (B.n * B.n))
        grad = torch.linalg.norm((2.0 * (C * vir[:, None, :]).transpose(1, 2) @ F_new @ Co).reshape(k, -1), dim=1)
        done = ((torch.abs(e_new - energy) < conv_tol) & (rms < conv_tol_density)
                & (grad < conv_tol_grad) & (cycle > 0))
        energy = torch.where(active, e_new, energy)
        D = D_new
        F = torch.where(active[:, None, None], F_new, F)
        was = active
        active = active & ~done
        # one host synchronisation per iteration: which molecules are still running
        flags = torch.stack((was, active)).cpu().numpy()
        was_h, active_h = flags[0], flags[1]
        cycles[was_h] = cycle + 1
        iteration_s.append(time.perf_counter() - ti)
        if not bool(active.any()):
            break
    torch.cuda.synchronize()
    t_end = time.perf_counter()
    eig.check()
    energies = energy.cpu().numpy()
    converged = ~active.cpu().numpy()
    d_final = torch.zeros_like(d_packed)
    B.scatter(d_final, D)
    groups = sorted({(int(n), int(o)) for n, o in zip(plan.mol_n, nelectrons // 2)})
    return BatchRHFResult(energies, converged, cycles, d_final, iteration_s,
                          {"integrals_s": t_int - t0, "scf_s": t_end - t_int, "jk_builds": n_jk,
                           "padded_order": N, "shapes": groups})
