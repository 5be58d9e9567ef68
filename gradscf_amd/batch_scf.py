"""Batched RHF over many independent molecules of uneven size.

Each molecule follows exactly the single-molecule driver in gradscf_amd.scf
(the reference semantics cited there: Loewdin orthogonaliser, core guess,
8-deep DIIS on C0^T (FDS - SDF) C0, incremental J/K on D_new - D_old,
convergence on |dE|, rms(dD) and the orbital gradient).  What changes is the
schedule:

  * one packed J/K call per SCF iteration covers every molecule (a batched
    plan: one launch per quartet class for the whole batch), optionally
    replayed from a captured HIP graph;
  * molecules with the same (AO count, occupied count) form a group whose
    diagonalisation, DIIS extrapolation and convergence tests run as batched
    FP64 tensor operations;
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
from .linalg import eigh

DIIS_SPACE = 8


@dataclass
class BatchRHFResult:
    energies: np.ndarray          # (nmol,) Eh
    converged: np.ndarray         # (nmol,) bool
    cycles: np.ndarray            # (nmol,) SCF cycles until convergence (or max_cycle)
    density: torch.Tensor         # packed (nn_total,) final densities
    iteration_s: list             # wall time of every SCF iteration (all molecules)
    timings: dict = field(default_factory=dict)


class _Group:
    """Molecules sharing (n, nocc): batched tensors (k, n, n) and gather indices."""

    def __init__(self, mols, n, nocc, offsets, device):
        self.mols = np.asarray(mols)
        self.k, self.n, self.nocc = len(mols), n, nocc
        base = torch.as_tensor(np.asarray(offsets)[self.mols], dtype=torch.int64, device=device)
        self.idx = (base[:, None] + torch.arange(n * n, device=device)[None, :]).reshape(-1)

    def gather(self, packed):
        return packed[self.idx].view(self.k, self.n, self.n)

    def scatter(self, packed, mats):
        packed[self.idx] = mats.reshape(-1)


def _orthogonalizer(s, eps=1e-10):
    w, v = eigh(s)
    return (v * torch.clamp(w, min=eps).rsqrt()[:, None, :]) @ v.transpose(1, 2)


def _diagonalize(f, x):
    fo = x.transpose(1, 2) @ f @ x
    fo = 0.5 * (fo + fo.transpose(1, 2))
    e, c = eigh(fo)
    return e, x @ c


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
    coeff = torch.linalg.solve(b, rhs)[:, :m] * valid
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
    enucs = enucs.to(dev)
    keys = {}
    for m in range(nmol):
        keys.setdefault((int(plan.mol_n[m]), int(nelectrons[m]) // 2), []).append(m)
    groups = [_Group(ms, n, nocc, plan.mol_offset, dev) for (n, nocc), ms in sorted(keys.items())]
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

    state = []
    d_packed = torch.zeros(plan.nn_total, dtype=torch.float64, device=dev)
    for g in groups:
        S, H = g.gather(s), g.gather(h)
        X = _orthogonalizer(S)
        _, C = _diagonalize(H, X)
        Co = C[:, :, :g.nocc]
        D = 2.0 * Co @ Co.transpose(1, 2)
        g.scatter(d_packed, D)
        state.append({"S": S, "H": H, "X": X, "C0": C, "D": D,
                      "fh": torch.zeros((g.k, DIIS_SPACE, g.n, g.n), dtype=torch.float64, device=dev),
                      "eh": torch.zeros((g.k, DIIS_SPACE, g.n * g.n), dtype=torch.float64, device=dev),
                      "gram": torch.zeros((g.k, DIIS_SPACE, DIIS_SPACE), dtype=torch.float64, device=dev),
                      "energy": torch.zeros(g.k, dtype=torch.float64, device=dev),
                      "active": torch.ones(g.k, dtype=torch.bool, device=dev),
                      "cycles": np.zeros(g.k, dtype=np.int64),
                      "enuc": enucs[torch.as_tensor(g.mols, device=dev)]})
    j_packed, k_packed = jk(d_packed)
    j_packed, k_packed = j_packed.clone(), k_packed.clone()
    for g, st in zip(groups, state):
        st["J"], st["K"] = g.gather(j_packed), g.gather(k_packed)
        st["F"] = st["H"] + st["J"] - 0.5 * st["K"]
    head = count = 0
    iteration_s = []
    n_jk = 1
    torch.cuda.synchronize()
    for cycle in range(max_cycle):
        ti = time.perf_counter()
        dd_packed = torch.zeros_like(d_packed)
        d_new_all = []
        # every group's DIIS step is queued before the first diagonalisation,
        # whose host transfer or error check synchronises with the device
        f_effs = []
        for g, st in zip(groups, state):
            F, D, S, C0 = st["F"], st["D"], st["S"], st["C0"]
            f_eff = F
            if cycle >= 1:
                err = C0.transpose(1, 2) @ (F @ D @ S - S @ D @ F) @ C0
                st["fh"][:, head] = F
                st["eh"][:, head] = err.reshape(g.k, -1)
                row = torch.bmm(st["eh"], st["eh"][:, head, :, None])[:, :, 0]
                st["gram"][:, head, :] = row
                st["gram"][:, :, head] = row
            if cycle >= 1 and min(count + 1, DIIS_SPACE) >= 2:
                f_eff = _diis_solve(st["fh"], st["gram"], min(count + 1, DIIS_SPACE))
            f_effs.append(f_eff)
        for (g, st), f_eff in zip(zip(groups, state), f_effs):
            D = st["D"]
            mo_e, C = _diagonalize(f_eff, st["X"])
            Co = C[:, :, :g.nocc]
            D_new = 2.0 * Co @ Co.transpose(1, 2)
            D_new = torch.where(st["active"][:, None, None], D_new, D)
            g.scatter(dd_packed, D_new - D)
            d_new_all.append((D_new, C))
        if cycle >= 1:
            head = (head + 1) % DIIS_SPACE
            count = min(count + 1, DIIS_SPACE)
        dj_packed, dk_packed = jk(dd_packed)
        n_jk += 1
        flags = []
        for (g, st), (D_new, C) in zip(zip(groups, state), d_new_all):
            st["J"] = st["J"] + g.gather(dj_packed)
            st["K"] = st["K"] + g.gather(dk_packed)
            F_new = st["H"] + st["J"] - 0.5 * st["K"]
            e_new = (D_new * (st["H"] + 0.5 * st["J"] - 0.25 * st["K"])).sum(dim=(1, 2)) + st["enuc"]
            rms = torch.sqrt(torch.mean((D_new - st["D"]).reshape(g.k, -1) ** 2, dim=1))
            grad = torch.linalg.norm((2.0 * C[:, :, g.nocc:].transpose(1, 2) @ F_new @ C[:, :, :g.nocc])
                                     .reshape(g.k, -1), dim=1)
            done = ((torch.abs(e_new - st["energy"]) < conv_tol) & (rms < conv_tol_density)
                    & (grad < conv_tol_grad) & (cycle > 0))
            act = st["active"]
            st["energy"] = torch.where(act, e_new, st["energy"])
            st["D"] = D_new
            st["F"] = torch.where(act[:, None, None], F_new, st["F"])
            st["active"] = act & ~done
            flags.append((act, st["active"]))
        # one host synchronisation per iteration: which molecules are still running
        still = torch.cat([a for _, a in flags]).any()
        was = [a.cpu().numpy() for a, _ in flags]
        for (g, st), w in zip(zip(groups, state), was):
            st["cycles"][w] = cycle + 1
        iteration_s.append(time.perf_counter() - ti)
        if not bool(still):
            break
    torch.cuda.synchronize()
    t_end = time.perf_counter()
    energies = np.zeros(nmol)
    converged = np.zeros(nmol, dtype=bool)
    cycles = np.zeros(nmol, dtype=np.int64)
    d_final = torch.zeros_like(d_packed)
    for g, st in zip(groups, state):
        energies[g.mols] = st["energy"].cpu().numpy()
        converged[g.mols] = ~st["active"].cpu().numpy()
        cycles[g.mols] = st["cycles"]
        g.scatter(d_final, st["D"])
    return BatchRHFResult(energies, converged, cycles, d_final, iteration_s,
                          {"integrals_s": t_int - t0, "scf_s": t_end - t_int, "jk_builds": n_jk,
                           "groups": [(g.n, g.nocc, g.k) for g in groups]})
