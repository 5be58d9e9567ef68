"""RHF nuclear gradient on the MI355X engine.

dE/dR = sum D d(T+V)/dR - W dS/dR            (one-electron kernel)
      + d/dR [ 1/2 tr(D J(D)) - 1/4 tr(D K(D)) ]   (J/K coordinate VJP with
                                                    cot_J = D/2, cot_K = -D/4)
      + dE_nn/dR,
with W = 2 C_occ diag(e_occ) C_occ^T, the standard variational RHF gradient
the reference differentiates through its SCF (its coordinate VJP of direct J/K,
src/gradscf/integrals/backends/native/autodiff/direct_jk.py:14-108, and the
int1e geometry derivatives of csrc/geometry.cc).  The packed tables give every
shell its own basis-centre atm row; gradients on those rows are folded onto
the nucleus the shell sits on.  The J/K VJP needs an unscreened plan (the
reference's "Screened direct J/K geometry AD requires cutoff=0").
"""
from __future__ import annotations

import numpy as np
import torch

from .basis import normalized_shell_coefficients_torch
from .jk import DirectJKPlan


def shell_nuclei(centers, nuclear_coords) -> np.ndarray:
    """Nucleus index of every shell (the nucleus its centre coincides with)."""
    c = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    r = np.asarray(nuclear_coords, dtype=np.float64).reshape(-1, 3)
    d = np.linalg.norm(c[:, None, :] - r[None, :, :], axis=-1)
    idx = np.argmin(d, axis=1)
    if np.any(d[np.arange(len(c)), idx] > 1e-10):
        raise ValueError("every shell centre must sit on a nucleus")
    return idx


def nuclear_repulsion_gradient(charges, coords) -> np.ndarray:
    z = np.asarray(charges, dtype=np.float64)
    r = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    g = np.zeros_like(r)
    for a in range(len(z)):
        for b in range(len(z)):
            if a != b:
                dr = r[a] - r[b]
                g[a] -= z[a] * z[b] * dr / np.linalg.norm(dr) ** 3
    return g


def fold_rows(row_grad, nnuc: int, shell_atom) -> np.ndarray:
    """(natm_rows, 3) gradient over atm rows -> (nnuc, 3) over nuclei."""
    rg = np.asarray(row_grad, dtype=np.float64)
    g = rg[:nnuc].copy()
    np.add.at(g, np.asarray(shell_atom), rg[nnuc:nnuc + len(shell_atom)])
    return g


def rhf_gradient(plan: DirectJKPlan, density, mo_coeff, mo_energy, nocc: int, charges, coords,
                 shell_atom) -> np.ndarray:
    """Nuclear gradient (nnuc, 3) in Eh/bohr from a converged RHF solution.

    `plan` must be unscreened (cutoff 0) and built from the same tables as the
    SCF; density/mo_coeff/mo_energy are the SCF's (GPU tensors)."""
    if plan.cutoff > 0:
        raise ValueError("Screened direct J/K geometry AD requires cutoff=0")
    d = density.to(torch.float64).contiguous()
    co = mo_coeff[:, :nocc]
    w = 2.0 * (co * mo_energy[:nocc]) @ co.T
    g1 = plan.int1e_gradient(d, w.contiguous())
    g2 = plan.jk_geometry_vjp(d, 0.5 * d, -0.25 * d)
    rows = (g1 + g2).cpu().numpy()
    nnuc = len(charges)
    return fold_rows(rows, nnuc, shell_atom) + nuclear_repulsion_gradient(charges, coords)


def env_to_basis_gradient(env_grad, bas, topology, parameters):
    """Chain an env-slot gradient (exponent slots and normalised-coefficient
    slots) through the basis normalisation: returns (dE/d exponents,
    dE/d raw coefficients) per shell, shaped like parameters.exponents and
    parameters.coefficients (the reference's JAX normalisation chain,
    src/gradscf/integrals/basis/normalization.py:57-63)."""
    g = np.asarray(env_grad, dtype=np.float64)
    d_exp, d_coef = [], []
    for s, (l, a, c) in enumerate(zip(topology.angular_momenta, parameters.exponents, parameters.coefficients)):
        row = bas[s]
        npr, nct = int(row[2]), int(row[3])
        ge = torch.as_tensor(g[int(row[5]):int(row[5]) + npr])
        gc = torch.as_tensor(g[int(row[6]):int(row[6]) + npr * nct].reshape(nct, npr).T.copy())
        at = torch.tensor(np.asarray(a, dtype=np.float64), requires_grad=True)
        ct = torch.tensor(np.asarray(c, dtype=np.float64).reshape(npr, nct), requires_grad=True)
        e = normalized_shell_coefficients_torch(int(l), at, ct)
        loss = (gc * e).sum() + (ge * at).sum()
        loss.backward()
        d_exp.append(at.grad.numpy().copy())
        d_coef.append(ct.grad.numpy().reshape(np.asarray(c).shape).copy())
    return d_exp, d_coef


def rhf_basis_gradient(plan: DirectJKPlan, density, mo_coeff, mo_energy, nocc: int, bas, topology, parameters):
    """dE/d(exponents) and dE/d(raw contraction coefficients) of a converged
    RHF energy (variational: D and W held fixed), per shell.  `plan` must be
    unscreened and built from the packed tables of (topology, parameters)."""
    if plan.cutoff > 0:
        raise ValueError("Screened direct J/K basis-parameter AD requires cutoff=0")
    d = density.to(torch.float64).contiguous()
    co = mo_coeff[:, :nocc]
    w = (2.0 * (co * mo_energy[:nocc]) @ co.T).contiguous()
    env_grad = plan.int1e_param_vjp(d, w) + plan.jk_param_vjp(d, 0.5 * d, -0.25 * d)
    return env_to_basis_gradient(env_grad.cpu().numpy(), bas, topology, parameters)
