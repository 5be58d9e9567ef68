"""Dense FP64 eigen-solver placement for the SCF loops.

The SCF diagonalisations are small, and the GPU solver (rocSOLVER through
torch.linalg.eigh) is launch-bound there.  Measured on MI355X against host
LAPACK including the device<->host copies (ms per call):

  one matrix      n = 13: 0.42 vs 0.04    n = 384: 9.1 vs 7.2
                  n = 512: 12.0 vs 10.8   n = 608: 14.0 vs 16.5
  32 matrices     n = 19: 1.6 vs 0.6      n = 57: 11.4 vs 4.3
                  n = 64: 14.2            n = 65: 1.9    n = 76: 1.5 vs 11.0
                  n = 152: 3.7 vs 41.4
  256 matrices    n = 152: 9.1

The batched GPU path changes algorithm above n = 64, and the larger one is
the fast one, so a batch of matrices with n <= 64 is solved as matrices of
order 65: the padding block is a multiple of the identity placed above every
eigenvalue (a Gershgorin bound), which leaves the leading n eigenpairs those
of the original matrices (1.7-1.9 ms for 32 matrices of order 19..57).  Single
matrices up to n = 560 go to host LAPACK.  The arithmetic is FP64 throughout.
"""
from __future__ import annotations

import torch

HOST_MAX_SINGLE = 560   # one matrix: host LAPACK up to this dimension
PAD_BATCHED_TO = 65     # several matrices: smallest order of the fast batched GPU path


def on_host(shape) -> bool:
    """Whether a symmetric eigen-problem of this (batched) shape runs on the host."""
    n = shape[-1]
    batch = 1
    for s in shape[:-2]:
        batch *= s
    return batch == 1 and n <= HOST_MAX_SINGLE


def gershgorin_bound(a: torch.Tensor) -> torch.Tensor:
    """max over the batch of max_i sum_j |a_ij|: no eigenvalue exceeds it (scalar tensor)."""
    return a.abs().sum(dim=-1).amax()


def eigh(a: torch.Tensor):
    """torch.linalg.eigh of a symmetric (batched) FP64 matrix on the faster of
    the GPU solver and host LAPACK for its shape; results on a's device."""
    if not a.is_cuda:
        return torch.linalg.eigh(a)
    if on_host(a.shape):
        e, v = torch.linalg.eigh(a.cpu())
        return e.to(a.device), v.to(a.device)
    if a.dim() > 2 and a.shape[-1] < PAD_BATCHED_TO:
        return eigh_padded(a, PAD_BATCHED_TO)
    return torch.linalg.eigh(a)


def eigh_padded(a: torch.Tensor, order: int):
    """Eigenpairs of symmetric matrices a (..., n, n) solved as matrices of
    order `order` >= n: a in the leading block, the rest a multiple of the
    identity above every eigenvalue, so the leading n eigenpairs are a's."""
    n = a.shape[-1]
    lead = a.shape[:-2]
    p = torch.zeros(lead + (order, order), dtype=a.dtype, device=a.device)
    p[..., :n, :n] = a
    idx = torch.arange(n, order, device=a.device)
    p[..., idx, idx] = gershgorin_bound(a) + 1.0
    e, v = torch.linalg.eigh(p)
    return e[..., :n].contiguous(), v[..., :n, :n].contiguous()
