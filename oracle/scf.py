"""Oracle RHF (test infrastructure only): numpy restatement of the reference SCF.

Follows
  * Loewdin X = S^{-1/2} with eigenvalue clip eps=1e-10:
    src/gradscf/solvers/eigen/spectral.py:38-42
  * core-Hamiltonian guess, D = C occ C^T (occ = 2): src/gradscf/scf/rks.py:1121-1152,
    src/gradscf/scf/core.py:53-70
  * cycle body: DIIS from cycle >= 1 (push, extrapolate once 2 entries), error
    C0^T (F D S - S D F) C0 with C0 the core-guess MOs; diagonalise; incremental
    J/K on D_new - D_old; F = H + J - K/2; E = tr(DH) + tr(DJ)/2 - tr(DK)/4 + E_nn;
    convergence |dE| < conv_tol, rms(dD) < 1e-8, ||2 Cv^T F Co|| < 1e-7, cycle > 0:
    src/gradscf/scf/rks.py:166-179, 309-322, 684-709, 787-846, 848-1079
  * DIIS Gram solve scaled by max|B| with 50*eps diagonal regularisation, 8-deep ring:
    src/gradscf/solvers/nonlinear/diis.py:14-103
  * energy / Fock algebra: src/gradscf/scf/energy.py:23-31
"""
from __future__ import annotations

import numpy as np


def orthogonalizer(s: np.ndarray, eps: float = 1e-10) -> np.ndarray:
    w, v = np.linalg.eigh(s)
    return v @ np.diag(np.maximum(w, eps) ** -0.5) @ v.T


def diagonalize(f, x):
    fo = x.T @ f @ x
    fo = 0.5 * (fo + fo.T)
    e, c = np.linalg.eigh(fo)
    return e, x @ c


def diis_solve(fock_hist, err_hist, count):
    m = fock_hist.shape[0]
    valid = (np.arange(m) < count).astype(float)
    gram = err_hist @ err_hist.T
    scale = np.max(np.abs(gram))
    scale = scale if scale > 0 else 1.0
    gram = gram / scale
    top = gram * (valid[:, None] * valid[None, :]) + np.diag(valid * 50 * np.finfo(float).eps + (1 - valid))
    b = np.zeros((m + 1, m + 1))
    b[:m, :m] = top
    b[:m, m] = -valid
    b[m, :m] = -valid
    rhs = np.zeros(m + 1)
    rhs[m] = -1.0
    coeff = np.linalg.solve(b, rhs)[:m] * valid
    return np.tensordot(coeff, fock_hist, axes=(0, 0))


def rhf(h, s, nelectron, enuc, jk, *, conv_tol=1e-10, conv_tol_density=1e-8,
        conv_tol_grad=1e-7, max_cycle=80, diis_space=8):
    """jk(D) -> (J, K).  Returns dict(energy, converged, cycles, density, mo_coeff, mo_energy, J, K)."""
    n = h.shape[0]
    nocc = nelectron // 2
    occ = np.zeros(n)
    occ[:nocc] = 2.0
    x = orthogonalizer(s)
    mo_e, c = diagonalize(h, x)
    c0 = c.copy()
    d = (c * occ) @ c.T
    j, k = jk(d)
    fock = h + j - 0.5 * k
    energy = 0.0
    fh = np.zeros((diis_space, n, n))
    eh = np.zeros((diis_space, n * n))
    head = count = 0
    converged = False
    cycles = 0
    for cycle in range(max_cycle):
        fock_eff = fock
        if cycle >= 1:
            err = c0.T @ (fock @ d @ s - s @ d @ fock) @ c0
            fh[head] = fock
            eh[head] = err.ravel()
            head = (head + 1) % diis_space
            count = min(count + 1, diis_space)
            if count >= 2:
                fock_eff = diis_solve(fh, eh, count)
        mo_e, c = diagonalize(fock_eff, x)
        d_new = (c * occ) @ c.T
        dj, dk = jk(d_new - d)
        j, k = j + dj, k + dk
        fock_new = h + j - 0.5 * k
        e_new = (np.sum(d_new * h) + 0.5 * np.sum(d_new * j) - 0.25 * np.sum(d_new * k) + enuc)
        rms = np.sqrt(np.mean((d_new - d) ** 2))
        grad = np.linalg.norm(2.0 * c[:, nocc:].T @ fock_new @ c[:, :nocc])
        done = (abs(e_new - energy) < conv_tol and rms < conv_tol_density and grad < conv_tol_grad
                and cycle > 0)
        d, fock, energy = d_new, fock_new, e_new
        cycles = cycle + 1
        if done:
            converged = True
            break
    return dict(energy=energy, converged=converged, cycles=cycles, density=d, mo_coeff=c,
                mo_energy=mo_e, J=j, K=k, fock=fock)
