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

import numpy as np
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


class _RocSolver:
    """rocSOLVER's batched divide-and-conquer eigen-solver (the routine
    torch.linalg.eigh runs) called directly: no host synchronisation on its
    error codes, and one rocBLAS handle per stream so that batches of
    different order run concurrently."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            import ctypes
            import os
            root = os.environ.get("ROCM_PATH", "/opt/rocm")
            blas = ctypes.CDLL(os.path.join(root, "lib", "librocblas.so"))
            sol = ctypes.CDLL(os.path.join(root, "lib", "librocsolver.so"))
            P, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
            blas.rocblas_create_handle.argtypes = [ctypes.POINTER(P)]
            blas.rocblas_set_stream.argtypes = [P, P]
            blas.rocblas_destroy_handle.argtypes = [P]
            sol.rocsolver_dsyevd_strided_batched.argtypes = [P, i32, i32, i32, P, i32, i64, P, i64, P, i64, P, i32]
            cls._lib = (blas, sol)
        return cls._lib


class BatchedEigh:
    """Eigenpairs of the molecules' blocks of a padded (k, N, N) batch
    (molecule m in rows and columns < n_m), grouped by order: each group's
    blocks are solved at their own order (no padding work), the groups on
    separate streams (default priority), largest first.  Returns eigenvalues
    (k, N) ascending within each block (padding entries +inf) and
    eigenvectors (k, N, N) with each block's eigenvectors in its leading
    columns and the identity in the padding, the layout a padded solve with a
    lifted padding diagonal gives.  Molecules outside `active` are skipped
    (their rows of the result are left as they were).  Solver failures are
    recorded on the device; check() raises if any occurred.  Host tensors
    (tests of the layout) are solved group by group by torch.linalg.eigh."""

    ROCBLAS_EVECT_ORIGINAL = 211
    ROCBLAS_FILL_LOWER = 122

    def __init__(self, mol_n, device, max_streams: int = 8):
        import ctypes
        n = np.asarray(mol_n, dtype=np.int64)
        self.k, self.N = len(n), int(n.max())
        self.device = device
        self.groups = []  # (order, host indices)
        for order in sorted(set(n.tolist()), reverse=True):
            self.groups.append((order, np.nonzero(n == order)[0]))
        self.info = torch.zeros(len(self.groups), dtype=torch.int32, device=device)
        self._idx = {}
        self.streams, self.handles = [], []
        if torch.device(device).type != "cuda":
            return
        self.blas, self.sol = _RocSolver.lib()
        self.streams = [torch.cuda.Stream(device=device) for _ in range(min(max_streams, len(self.groups)))]
        self.handles = []
        for s in self.streams:
            h = ctypes.c_void_p()
            if self.blas.rocblas_create_handle(ctypes.byref(h)) != 0:
                raise RuntimeError("rocblas_create_handle failed")
            self.handles.append(h)
            if self.blas.rocblas_set_stream(h, ctypes.c_void_p(s.cuda_stream)) != 0:
                raise RuntimeError("rocblas_set_stream failed")

    def __del__(self):
        for h in getattr(self, "handles", []):
            self.blas.rocblas_destroy_handle(h)

    def _indices(self, g, active_h):
        order, idx = self.groups[g]
        if active_h is not None:
            idx = idx[active_h[idx]]
        key = (g, idx.tobytes())
        if key not in self._idx:
            self._idx = {kk: v for kk, v in self._idx.items() if kk[0] != g}
            self._idx[key] = torch.as_tensor(idx, device=self.device)
        return self._idx[key]

    def __call__(self, a: torch.Tensor, e_out=None, c_out=None, active_h=None):
        import ctypes
        k, N = self.k, self.N
        if a.shape != (k, N, N) or a.dtype != torch.float64:
            raise ValueError("BatchedEigh expects a (k, N, N) float64 batch")
        if e_out is None:
            e_out = torch.full((k, N), float("inf"), dtype=torch.float64, device=a.device)
        if c_out is None:
            c_out = torch.eye(N, dtype=torch.float64, device=a.device).repeat(k, 1, 1)
        if not a.is_cuda:
            for g, (order, _) in enumerate(self.groups):
                idx = self._indices(g, active_h)
                if idx.numel():
                    w, v = torch.linalg.eigh(a.index_select(0, idx)[:, :order, :order])
                    c_out[idx, :order, :order] = v
                    e_out[idx, :order] = w
            return e_out, c_out
        cur = torch.cuda.current_stream(a.device)
        fork = torch.cuda.Event()
        fork.record(cur)
        P = ctypes.c_void_p
        used = set()
        for g, (order, _) in enumerate(self.groups):
            idx = self._indices(g, active_h)
            if idx.numel() == 0:
                continue
            j = g % len(self.streams)
            s = self.streams[j]
            used.add(j)
            s.wait_event(fork)
            with torch.cuda.stream(s):
                blk = a.index_select(0, idx)[:, :order, :order].contiguous()
                w = torch.empty((idx.numel(), order), dtype=torch.float64, device=a.device)
                ework = torch.empty_like(w)
                info = torch.zeros(idx.numel(), dtype=torch.int32, device=a.device)
                st = self.sol.rocsolver_dsyevd_strided_batched(
                    self.handles[j], self.ROCBLAS_EVECT_ORIGINAL, self.ROCBLAS_FILL_LOWER, order,
                    P(blk.data_ptr()), order, order * order, P(w.data_ptr()), order, P(ework.data_ptr()), order,
                    P(info.data_ptr()), idx.numel())
                if st != 0:
                    raise RuntimeError(f"rocsolver_dsyevd_strided_batched failed (status {st})")
                # column-major eigenvectors: row i of the row-major buffer is eigenvector i
                c_out[idx, :order, :order] = blk.transpose(1, 2)
                e_out[idx, :order] = w
                self.info[g:g + 1] += info.abs().amax().reshape(1)
                for t in (blk, w, ework, info):
                    t.record_stream(s)
        for j in used:
            cur.wait_stream(self.streams[j])
        return e_out, c_out

    def check(self):
        if int(self.info.abs().max()) != 0:
            raise RuntimeError("the batched eigen-solver did not converge for some molecule")
