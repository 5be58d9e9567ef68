"""libcint-layout ATM/BAS/ENV tables.

Reference semantics followed (src/gradscf/integrals/backends/native/packing.py):
  * build_tables (:78-95): nuclear atoms first (charge Z), then one zero-charge
    basis-centre atom per shell; every atom row points at 3 ENV coordinate
    slots starting at ENV offset 20; bas row = (atom, l, nprim, nctr, kappa=0,
    ptr_exp, ptr_coeff, 0); coefficients stored column-major (ctr-major).
  * pack / pack_environment (:17-47): coordinates then exponents and normalised
    coefficients written into ENV.
"""
from __future__ import annotations

import numpy as np

from .basis import (BasisParameters, BasisTopology, normalized_shell_coefficients,
                    normalized_shell_coefficients_torch)

# libcint slot constants (ATM_SLOTS, BAS_SLOTS and ENV pointers)
ATM_SLOTS = 6
BAS_SLOTS = 8
CHARGE_OF, PTR_COORD, NUC_MOD_OF = 0, 1, 2
ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, KAPPA_OF, PTR_EXP, PTR_COEFF = 0, 1, 2, 3, 4, 5, 6
PTR_EXPCUTOFF = 0
PTR_RANGE_OMEGA = 8
PTR_ENV_START = 20


def build_tables(topology: BasisTopology):
    nnuc = len(topology.nuclear_charges)
    nsh = len(topology.angular_momenta)
    atm = np.zeros((nnuc + nsh, ATM_SLOTS), dtype=np.int32)
    atm[:nnuc, CHARGE_OF] = topology.nuclear_charges
    atm[:, NUC_MOD_OF] = 1
    atm[:, PTR_COORD] = PTR_ENV_START + np.arange(nnuc + nsh) * 3
    offset = PTR_ENV_START + 3 * (nnuc + nsh)
    bas = np.zeros((nsh, BAS_SLOTS), dtype=np.int32)
    for i, (l, np_, nc) in enumerate(zip(topology.angular_momenta, topology.primitive_counts,
                                        topology.contraction_counts)):
        bas[i, :4] = (nnuc + i, l, np_, nc)
        bas[i, PTR_EXP] = offset
        bas[i, PTR_COEFF] = offset + np_
        offset += np_ + np_ * nc
    return atm, bas, offset


def pack_environment(topology, parameters: BasisParameters, coefficients, *, origin=None):
    atm, bas, env_size = build_tables(topology)
    env = np.zeros(env_size, dtype=np.float64)
    if origin is not None:
        env[1:4] = np.asarray(origin, dtype=np.float64)
    coords = np.concatenate([np.asarray(parameters.nuclear_coords).reshape(-1, 3),
                             np.asarray(parameters.centers).reshape(-1, 3)], axis=0)
    for i in range(len(atm)):
        off = int(atm[i, PTR_COORD])
        env[off:off + 3] = coords[i]
    for i, (a, c) in enumerate(zip(parameters.exponents, coefficients)):
        ae, ce = int(bas[i, PTR_EXP]), int(bas[i, PTR_COEFF])
        a = np.asarray(a, dtype=np.float64)
        c = np.asarray(c, dtype=np.float64).reshape(a.size, -1)
        env[ae:ae + a.size] = a
        env[ce:ce + c.size] = c.T.ravel()
    return atm, bas, env


def pack(topology: BasisTopology, parameters: BasisParameters, *, origin=None):
    """(atm int32[natm,6], bas int32[nbas,8], env float64[nenv]) with normalised coefficients."""
    coefficients = tuple(normalized_shell_coefficients(l, a, c) for l, a, c in
                         zip(topology.angular_momenta, parameters.exponents, parameters.coefficients))
    return pack_environment(topology, parameters, coefficients, origin=origin)


def pack_torch(topology: BasisTopology, parameters: BasisParameters, *, device=None):
    """(atm, bas, env) like pack, with env a float64 torch tensor assembled by
    differentiable operations from parameters whose exponents, coefficients,
    centres or nuclear coordinates may be tensors that require gradients."""
    import torch
    atm, bas, env_size = build_tables(topology)

    def t(x):
        return torch.as_tensor(x, dtype=torch.float64, device=device).reshape(-1)

    pieces = [torch.zeros(PTR_ENV_START, dtype=torch.float64, device=device),
              t(parameters.nuclear_coords), t(parameters.centers)]
    for l, a, c in zip(topology.angular_momenta, parameters.exponents, parameters.coefficients):
        a = t(a)
        c = torch.as_tensor(c, dtype=torch.float64, device=a.device).reshape(a.numel(), -1)
        pieces.append(a)
        pieces.append(normalized_shell_coefficients_torch(int(l), a, c).T.reshape(-1))
    env = torch.cat([p.to(pieces[0].device) for p in pieces])
    if env.numel() != env_size:
        raise ValueError("basis parameters do not match the topology")
    return atm, bas, env


def ao_offsets(bas: np.ndarray, cart: bool) -> np.ndarray:
    """ao[s+1] = ao[s] + nctr * (cart ? (l+1)(l+2)/2 : 2l+1)."""
    ao = np.zeros(len(bas) + 1, dtype=np.int64)
    for s, row in enumerate(bas):
        l, nc = int(row[ANG_OF]), int(row[NCTR_OF])
        ao[s + 1] = ao[s] + nc * ((l + 1) * (l + 2) // 2 if cart else 2 * l + 1)
    return ao
