"""Dense FP64 eigen-solver placement for the SCF loops.

The SCF diagonalisations are small, and the GPU solver (rocSOLVER through
torch.linalg.eigh) is launch-bound there.  Measured on MI355X against host
LAPACK including the device<->host copies (ms per call):

  one matrix      n = 13: 0.42 vs 0.04    n = 384: 9.1 vs 7.2
                  n = 512: 12.0 vs 10.8   n = 608: 14.0 vs 16.5
  32 matrices     n = 19: 1.5 vs 0.7      n = 57: 11.4 vs 5.2
                  n = 76: 1.5 vs 7.9      n = 152: 3.6 vs 46.3

(the batched GPU path changes algorithm above n = 64).  `eigh` follows these
crossovers; the arithmetic is FP64 either way.
"""
from __future__ import annotations

import torch

HOST_MAX_SINGLE = 560   # one matrix: host LAPACK up to this dimension
HOST_MAX_BATCHED = 64   # several matrices: host LAPACK up to this dimension


def on_host(shape) -> bool:
    """Whether a symmetric eigen-problem of this (batched) shape runs on the host."""
    n = shape[-1]
    batch = 1
    for s in shape[:-2]:
        batch *= s
    return n <= (HOST_MAX_SINGLE if batch == 1 else HOST_MAX_BATCHED)


def eigh(a: torch.Tensor):
    """torch.linalg.eigh of a symmetric (batched) FP64 matrix on the faster of
    the GPU solver and host LAPACK for its shape; results on a's device."""
    if not a.is_cuda or not on_host(a.shape):
        return torch.linalg.eigh(a)
    e, v = torch.linalg.eigh(a.cpu())
    return e.to(a.device), v.to(a.device)
