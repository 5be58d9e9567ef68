"""Direct J/K on the MI355X engine, behind the reference's Python call surface.

Mirrors:
  * direct_jk(atm, bas, env, density, *, cart, cutoff, workers, with_j, with_k,
    coordinates) -- src/gradscf/integrals/backends/native/jk.py:399-446
    (argument checks, complex densities split into real/imaginary parts,
    (J, K) returned with the density's shape, zeros for an unselected operator);
  * IntegralPlan.get_jk / make_plan -- src/gradscf/integrals/plan.py:71-92,184-190;
  * NativeDirectBasis.get_jk -- src/gradscf/integrals/backends/native/jk.py:98-107.

Densities and results are torch tensors on the GPU (float64).  Device plans
(pairs, primitive-pair data, Schwarz bounds) are cached per (tables, cutoff)
and reused across SCF iterations; the reference rebuilds them on every call.
"""
from __future__ import annotations

import ctypes
from collections import OrderedDict
from dataclasses import dataclass
from numbers import Integral

import numpy as np
import torch

from . import _native
from .basis import BasisParameters, BasisTopology
from .packing import PTR_COORD, pack, pack_torch

_PLAN_CACHE: "OrderedDict[tuple, DirectJKPlan]" = OrderedDict()
_PLAN_CACHE_SIZE = 8


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("The MI355X J/K engine needs a GPU (torch.cuda is not available).")
    return torch.device("cuda", torch.cuda.current_device())


def _stream_handle() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def check_out_buffer(out: torch.Tensor, shape: tuple, device: torch.device) -> None:
    """A caller-supplied result buffer must be exactly what the native build writes
    through its pointer: float64, contiguous, on the density's GPU, of `shape`."""
    if not isinstance(out, torch.Tensor):
        raise ValueError("Direct J/K output must be a torch tensor")
    if out.dtype != torch.float64 or not out.is_cuda or out.device != device:
        raise ValueError("Direct J/K output must be float64 on the density's GPU")
    if tuple(out.shape) != tuple(shape) or not out.is_contiguous():
        raise ValueError("Direct J/K output shape mismatch")


def concat_tables(tables):
    """Concatenate per-molecule libcint tables [(atm, bas, env), ...] into one
    (atm, bas, env) plus the molecule offsets (mol_atm, mol_bas) of a batched
    plan: env pointers and atom indices are shifted by the preceding sizes."""
    atms, bass, envs = [], [], []
    mol_atm, mol_bas = [0], [0]
    env_off = 0
    for atm, bas, env in tables:
        atm = np.array(atm, dtype=np.int32, copy=True)
        bas = np.array(bas, dtype=np.int32, copy=True)
        env = np.asarray(env, dtype=np.float64)
        atm[:, 1] += env_off            # PTR_COORD
        atm[:, 4] = np.where(atm[:, 4] > 0, atm[:, 4] + env_off, atm[:, 4])  # PTR_ZETA when set
        bas[:, 0] += mol_atm[-1]        # ATOM_OF
        bas[:, 5] += env_off            # PTR_EXP
        bas[:, 6] += env_off            # PTR_COEFF
        atms.append(atm)
        bass.append(bas)
        envs.append(env)
        env_off += len(env)
        mol_atm.append(mol_atm[-1] + len(atm))
        mol_bas.append(mol_bas[-1] + len(bas))
    return (np.concatenate(atms), np.concatenate(bass), np.concatenate(envs),
            np.asarray(mol_atm, dtype=np.int32), np.asarray(mol_bas, dtype=np.int32))


class DirectJKPlan:
    """Owns one device plan (gscf_plan) for fixed tables and Schwarz cutoff.

    With mol_atm / mol_bas (offsets of nmol molecules in concatenated tables,
    see concat_tables) the plan is batched: every molecule is screened and
    contracted on its own, and densities / results are packed, molecule m's
    n_m x n_m matrix starting at element mol_offset[m] of a density of
    nn_total elements (pack / unpack convert)."""

    def __init__(self, atm, bas, env, *, cart: bool, cutoff: float = 0.0, mol_atm=None, mol_bas=None):
        L = _native.lib()
        self.atm = np.ascontiguousarray(atm, dtype=np.int32)
        self.bas = np.ascontiguousarray(bas, dtype=np.int32)
        self.env = np.ascontiguousarray(env, dtype=np.float64)
        if self.atm.ndim != 2 or self.atm.shape[1] != 6 or self.bas.ndim != 2 or self.bas.shape[1] != 8:
            raise ValueError("Invalid libcint table dimensions or operator")
        self.cart = bool(cart)
        self.cutoff = float(cutoff)
        _device()
        handle = ctypes.c_void_p()
        self.mol_atm = np.asarray([0, len(self.atm)] if mol_atm is None else mol_atm, dtype=np.int32)
        self.mol_bas = np.asarray([0, len(self.bas)] if mol_bas is None else mol_bas, dtype=np.int32)
        if mol_atm is None and mol_bas is None:
            _native.check(L.gscf_plan_create(
                self.atm.ctypes.data, len(self.atm), self.bas.ctypes.data, len(self.bas),
                self.env.ctypes.data, len(self.env), int(self.cart), ctypes.c_double(self.cutoff),
                _stream_handle(), ctypes.byref(handle)))
        else:
            ma = np.ascontiguousarray(mol_atm, dtype=np.int32)
            mb = np.ascontiguousarray(mol_bas, dtype=np.int32)
            if ma.ndim != 1 or ma.shape != mb.shape or len(ma) < 2:
                raise ValueError("A batched plan needs per-molecule atm and bas offsets")
            _native.check(L.gscf_plan_create_batch(
                self.atm.ctypes.data, len(self.atm), self.bas.ctypes.data, len(self.bas),
                self.env.ctypes.data, len(self.env), len(ma) - 1, ma.ctypes.data, mb.ctypes.data,
                int(self.cart), ctypes.c_double(self.cutoff), _stream_handle(), ctypes.byref(handle)))
        self._handle = handle
        self.incore_enabled = False
        info = _native.PlanInfo()
        _native.check(L.gscf_plan_get_info(handle, ctypes.byref(info)))
        self.info = {k: getattr(info, k) for k, _ in _native.PlanInfo._fields_}
        self.nao = int(info.n)
        self.nmol = int(info.nmol)
        self.nn_total = int(info.nn_total)
        self.mol_n = np.zeros(self.nmol, dtype=np.int32)
        self.mol_offset = np.zeros(self.nmol, dtype=np.int64)
        _native.check(L.gscf_plan_get_molecules(handle, self.nmol, self.mol_n.ctypes.data,
                                                self.mol_offset.ctypes.data))

    @classmethod
    def batch(cls, tables, *, cart: bool, cutoff: float = 0.0) -> "DirectJKPlan":
        """Batched plan over per-molecule tables [(atm, bas, env), ...]."""
        atm, bas, env, mol_atm, mol_bas = concat_tables(tables)
        return cls(atm, bas, env, cart=cart, cutoff=cutoff, mol_atm=mol_atm, mol_bas=mol_bas)

    def pack(self, mats) -> torch.Tensor:
        """Per-molecule (..., n_m, n_m) tensors -> packed (..., nn_total) tensor."""
        if len(mats) != self.nmol:
            raise ValueError("one matrix per molecule is required")
        lead = tuple(mats[0].shape[:-2])
        return torch.cat([m.reshape(*lead, -1) for m in mats], dim=-1).contiguous()

    def unpack(self, packed: torch.Tensor) -> list:
        """Packed (..., nn_total) tensor -> list of (..., n_m, n_m) views."""
        lead = tuple(packed.shape[:-1])
        return [packed[..., int(o):int(o) + int(n) * int(n)].reshape(*lead, int(n), int(n))
                for o, n in zip(self.mol_offset, self.mol_n)]

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _native.lib().gscf_plan_destroy(h)
            except Exception:
                pass
            self._handle = None

    def jk(self, density: torch.Tensor, *, with_j=True, with_k=True, symmetric=None,
           validate=True, out: torch.Tensor | None = None, mixed_threshold: float | None = None) -> torch.Tensor:
        """density (nb, n, n) float64 on the GPU -> out (nops, nb, n, n).

        mixed_threshold: evaluate in FP32 every quartet whose Q_ij Q_kl max|D|
        (max over the six density blocks it touches) is below this value;
        roots/weights, prefactors and all J/K accumulation stay FP64."""
        d = density
        n = self.nao
        if self.nmol == 1:
            if d.dim() == 2:
                d = d[None]
            ok = d.dim() == 3 and tuple(d.shape[1:]) == (n, n)
            tail = (n, n)
        else:
            if d.dim() == 1:
                d = d[None]
            ok = d.dim() == 2 and d.shape[1] == self.nn_total
            tail = (self.nn_total,)
        if not ok or d.dtype != torch.float64 or not d.is_cuda:
            raise ValueError("Direct J/K shape mismatch")
        d = d.contiguous()
        nb = d.shape[0]
        nops = int(bool(with_j)) + int(bool(with_k))
        if out is None:
            out = torch.empty((nops, nb) + tail, dtype=torch.float64, device=d.device)
        else:
            check_out_buffer(out, (nops, nb) + tail, d.device)
        if symmetric is None:
            if self.nmol == 1:
                symmetric = bool(torch.equal(d, d.transpose(1, 2)))
            else:
                symmetric = all(bool(torch.equal(m, m.transpose(-1, -2))) for m in self.unpack(d))
        flags = (_native.FLAG_SYMMETRIC_DENSITY if symmetric else 0) | (0 if validate else _native.FLAG_NO_VALIDATE)
        thr = 0.0
        if mixed_threshold is not None:
            flags |= _native.FLAG_MIXED_PRECISION
            thr = float(mixed_threshold)
        _native.check(_native.lib().gscf_jk_build(
            self._handle, d.data_ptr(), nb, out.data_ptr(), int(bool(with_j)), int(bool(with_k)),
            flags, ctypes.c_double(thr), _stream_handle()))
        return out

    def schwarz(self):
        """Parent-shell Schwarz bounds (nbas x nbas, lower triangle) of a screened
        plan; a list of per-molecule blocks for a batched plan."""
        if self.nmol == 1:
            q = np.zeros((len(self.bas), len(self.bas)))
            _native.check(_native.lib().gscf_plan_get_schwarz(self._handle, q.ctypes.data))
            return q
        counts = np.diff(self.mol_bas).astype(np.int64)
        flat = np.zeros(int(np.sum(counts ** 2)))
        _native.check(_native.lib().gscf_plan_get_schwarz(self._handle, flat.ctypes.data))
        out, o = [], 0
        for c in counts:
            out.append(flat[o:o + c * c].reshape(c, c))
            o += c * c
        return out

    def device_bytes(self) -> int:
        """Device memory the plan holds now (tables, workspace, in-core integrals)."""
        info = _native.PlanInfo()
        _native.check(_native.lib().gscf_plan_get_info(self._handle, ctypes.byref(info)))
        return int(info.device_bytes)

    def incore_bytes(self) -> int:
        """Device bytes that keeping the plan's integrals in core would take."""
        b = ctypes.c_int64()
        _native.check(_native.lib().gscf_plan_incore_bytes(self._handle, ctypes.byref(b)))
        return int(b.value)

    def incore(self, enable: bool = True) -> None:
        """Evaluate every screened-in quartet block once and keep it in device
        memory, so that later jk() calls digest stored integrals (the packed-ERI
        route of the reference, csrc/packed_eri.cc + csrc/packed_jk.cc); False
        releases them."""
        _native.check(_native.lib().gscf_plan_incore(self._handle, int(bool(enable)), _stream_handle()))
        self.incore_enabled = bool(enable)

    def profile(self, enable: bool) -> None:
        """Start (and reset) or stop per-class HIP-event timing of jk() launches."""
        _native.check(_native.lib().gscf_plan_profile(self._handle, int(bool(enable))))

    def class_times(self) -> list:
        """[(la, lb, lc, ld), summed_ms, launches] per quartet class since profile(True)."""
        cap = 64
        cls = np.zeros((cap, 4), dtype=np.int32)
        ms = np.zeros(cap)
        la = np.zeros(cap, dtype=np.int64)
        cnt = ctypes.c_int32()
        _native.check(_native.lib().gscf_plan_class_times(self._handle, cap, cls.ctypes.data, ms.ctypes.data,
                                                          la.ctypes.data, ctypes.byref(cnt)))
        return [(tuple(int(v) for v in cls[i]), float(ms[i]), int(la[i])) for i in range(cnt.value)]

    def jk_geometry_vjp(self, density, cot_j=None, cot_k=None) -> torch.Tensor:
        """d/dR [sum(cot_j * J(D)) + sum(cot_k * K(D))] for every atm-row coordinate,
        as (natm, 3); the reference's coordinate VJP of direct J/K (cutoff must be 0)."""
        dev = _device()
        n = self.nao

        def prep(x):
            if x is None:
                return None
            x = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x))
            x = x.to(dev, torch.float64).contiguous()
            if tuple(x.shape) != ((n, n) if self.nmol == 1 else (self.nn_total,)):
                raise ValueError("geometry VJP inputs must be (n, n) (packed for a batched plan)")
            return x

        d, cj, ck = prep(density), prep(cot_j), prep(cot_k)
        grad = torch.zeros((len(self.atm), 3), dtype=torch.float64, device=dev)
        _native.check(_native.lib().gscf_jk_geometry_vjp(
            self._handle, d.data_ptr(), None if cj is None else cj.data_ptr(),
            None if ck is None else ck.data_ptr(), grad.data_ptr(), _stream_handle()))
        return grad

    def int1e_gradient(self, density, wdensity) -> torch.Tensor:
        """sum D d(T+V)/dR - W dS/dR for every atm-row coordinate, as (natm, 3):
        basis-centre rows carry the basis-function derivatives, nuclear rows
        the Hellmann-Feynman term."""
        dev = _device()
        shape = (self.nao, self.nao) if self.nmol == 1 else (self.nn_total,)
        d = torch.as_tensor(density).to(dev, torch.float64).contiguous()
        w = torch.as_tensor(wdensity).to(dev, torch.float64).contiguous()
        if tuple(d.shape) != shape or tuple(w.shape) != shape:
            raise ValueError("int1e gradient inputs must be (n, n) (packed for a batched plan)")
        grad = torch.zeros((len(self.atm), 3), dtype=torch.float64, device=dev)
        _native.check(_native.lib().gscf_int1e_gradient(self._handle, d.data_ptr(), w.data_ptr(), grad.data_ptr(),
                                                        _stream_handle()))
        return grad

    def _env_vjp(self, fn, *mats) -> torch.Tensor:
        dev = _device()
        shape = (self.nao, self.nao) if self.nmol == 1 else (self.nn_total,)
        ptrs = []
        keep = []
        for m in mats:
            if m is None:
                ptrs.append(None)
                continue
            t = torch.as_tensor(m).to(dev, torch.float64).contiguous()
            if tuple(t.shape) != shape:
                raise ValueError("basis-parameter VJP inputs must be (n, n) (packed for a batched plan)")
            keep.append(t)
            ptrs.append(t.data_ptr())
        out = torch.zeros(len(self.env), dtype=torch.float64, device=dev)
        _native.check(fn(self._handle, *ptrs, out.data_ptr(), _stream_handle()))
        return out

    def jk_param_vjp(self, density, cot_j=None, cot_k=None) -> torch.Tensor:
        """d/d env [sum(cot_j * J(D)) + sum(cot_k * K(D))] at every exponent and
        (normalised) coefficient slot, as an env-sized vector (cutoff must be 0)."""
        return self._env_vjp(_native.lib().gscf_jk_param_vjp, density, cot_j, cot_k)

    def int1e_param_vjp(self, density, wdensity) -> torch.Tensor:
        """d/d env [sum D (T + V) - W S] at every exponent and coefficient slot."""
        return self._env_vjp(_native.lib().gscf_int1e_param_vjp, density, wdensity)

    def int1e(self):
        """(S, T, V) as (n, n) float64 GPU tensors in the plan's AO basis
        (packed (nn_total,) tensors for a batched plan)."""
        dev = _device()
        shape = (self.nao, self.nao) if self.nmol == 1 else (self.nn_total,)
        s, t, v = (torch.empty(shape, dtype=torch.float64, device=dev) for _ in range(3))
        _native.check(_native.lib().gscf_int1e(self._handle, s.data_ptr(), t.data_ptr(), v.data_ptr(),
                                               _stream_handle()))
        return s, t, v

    def eri_blocks(self, quartets) -> list:
        """Cartesian contracted blocks for parent-shell quartets, each as an array
        [d0, d1, d2, d3] (libcint layout transposed to C order)."""
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.int32).reshape(-1, 4))
        ao = [0]
        for row in self.bas:
            ao.append(ao[-1] + int(row[3]) * (int(row[1]) + 1) * (int(row[1]) + 2) // 2)
        dims = [[ao[s + 1] - ao[s] for s in quart] for quart in q]
        sizes = [int(np.prod(d)) for d in dims]
        offs = np.zeros(len(q), dtype=np.int64)
        if len(q):
            offs[1:] = np.cumsum(sizes)[:-1]
        out = torch.zeros(max(1, int(sum(sizes))), dtype=torch.float64, device=_device())
        _native.check(_native.lib().gscf_eri_blocks(self._handle, q.ctypes.data, len(q), offs.ctypes.data,
                                                    out.data_ptr(), _stream_handle()))
        h = out.cpu().numpy()
        return [h[o:o + s].reshape(d[3], d[2], d[1], d[0]).transpose(3, 2, 1, 0)
                for o, s, d in zip(offs, sizes, dims)]


def get_plan(atm, bas, env, *, cart: bool, cutoff: float) -> DirectJKPlan:
    atm = np.ascontiguousarray(atm, dtype=np.int32)
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    key = (atm.tobytes(), bas.tobytes(), env.tobytes(), bool(cart), float(cutoff))
    plan = _PLAN_CACHE.get(key)
    if plan is None:
        plan = DirectJKPlan(atm, bas, env, cart=cart, cutoff=cutoff)
        _PLAN_CACHE[key] = plan
        while len(_PLAN_CACHE) > _PLAN_CACHE_SIZE:
            _PLAN_CACHE.popitem(last=False)
    else:
        _PLAN_CACHE.move_to_end(key)
    return plan


def _validate_selection(with_j, with_k):
    if not all(isinstance(flag, (bool, np.bool_)) for flag in (with_j, with_k)):
        raise ValueError("with_j and with_k must be static booleans.")


def direct_jk(atm, bas, env, density, *, cart, cutoff=0.0, workers=1, with_j=True, with_k=True,
              coordinates=None):
    """Contract J/K for density (..., n, n); returns (J, K) shaped like density.

    Differentiable (torch.autograd, gradscf_amd.autodiff) in the density, in
    `coordinates` (every atm row, (natm, 3)) and in `env` when they are
    tensors that require gradients: density-linear and self-adjoint, the
    coordinate VJP and the env VJP (exponent, coefficient and coordinate
    slots); geometry and basis derivatives need cutoff 0."""
    from . import autodiff
    _validate_selection(with_j, with_k)
    if isinstance(workers, bool) or not isinstance(workers, Integral) or workers < 1:
        raise ValueError("workers must be a positive integer.")
    d = density if isinstance(density, torch.Tensor) else torch.as_tensor(np.asarray(density))
    if d.dim() < 2:
        raise ValueError("Direct J/K requires (...,n,n) float64/complex128 density.")
    n = d.shape[-1]
    if tuple(d.shape[-2:]) != (n, n) or d.dtype not in (torch.float64, torch.complex128):
        raise ValueError("Direct J/K requires (...,n,n) float64/complex128 density.")
    env_t = env if isinstance(env, torch.Tensor) else None
    env = env.detach().cpu().numpy() if env_t is not None else np.asarray(env)
    if env.dtype != np.float64:
        raise ValueError("Direct J/K requires float64 environment.")
    if not (with_j or with_k):
        return torch.zeros_like(d), torch.zeros_like(d)
    coords_t = coordinates if isinstance(coordinates, torch.Tensor) else None
    if coordinates is not None:
        coordinates = (coords_t.detach().cpu().numpy() if coords_t is not None
                       else np.asarray(coordinates, dtype=np.float64)).astype(np.float64)
        if coordinates.shape != (len(atm), 3):
            raise ValueError("Direct J/K coordinates must contain every nuclear and basis-center atom.")
        env = env.copy()
        for i, row in enumerate(np.asarray(atm)):
            env[int(row[PTR_COORD]):int(row[PTR_COORD]) + 3] = coordinates[i]
    dev = _device()
    flat = d.reshape(-1, n, n).to(dev)
    count = flat.shape[0]
    real = torch.cat((flat.real, flat.imag), 0) if flat.is_complex() else flat
    plan = get_plan(atm, bas, env, cart=cart, cutoff=cutoff)
    if plan.nao != n:
        raise ValueError("Direct J/K shape mismatch")
    if autodiff.needs_grad(real, env_t, coords_t):
        op = autodiff.JKOperator(plan, bool(with_j), bool(with_k))
        if env_t is not None and coords_t is not None:
            # the env's coordinate slots are replaced by `coordinates`
            keep = torch.ones(len(env), dtype=torch.float64, device=env_t.device)
            for row in np.asarray(atm):
                keep[int(row[PTR_COORD]):int(row[PTR_COORD]) + 3] = 0.0
            env_t = env_t * keep
        out = autodiff.apply_jk(op, real.contiguous(), coords_t, env_t)
    else:
        out = plan.jk(real.contiguous(), with_j=with_j, with_k=with_k)
    if flat.is_complex():
        out = torch.complex(out[:, :count], out[:, count:])
    out = out.reshape(int(with_j) + int(with_k), *d.shape)
    return (out[0] if with_j else torch.zeros_like(out[0]),
            out[-1] if with_k else torch.zeros_like(out[0]))


def packed_eri_shape(nao: int, ndim: int):
    """Shape of packed ERIs (molecular/eri.py:21-23): the s8 triangle
    (ndim 1) or the np x np pair matrix (ndim 2), np = nao (nao + 1) / 2."""
    npair = nao * (nao + 1) // 2
    return (npair * (npair + 1) // 2,) if ndim == 1 else (npair, npair)


def packed_eri(atm, bas, env, *, cart: bool, aosym: str = "s8") -> torch.Tensor:
    """Every unique (ab|cd) of one molecule, unscreened, on the GPU: the
    reference's compact integrals (csrc/packed_eri.cc), aosym "s8" (triangle,
    layout 1) or "s4" (pair matrix, layout 0)."""
    if aosym not in ("s8", "s4"):
        raise ValueError("aosym must be 's8' or 's4'")
    atm = np.ascontiguousarray(atm, dtype=np.int32)
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    if atm.ndim != 2 or atm.shape[1] != 6 or bas.ndim != 2 or bas.shape[1] != 8:
        raise ValueError("Invalid libcint table dimensions or operator")
    L = bas[:, 1]
    nctr = bas[:, 3]
    n = int(np.sum(nctr * (((L + 1) * (L + 2)) // 2 if cart else 2 * L + 1)))
    shape = packed_eri_shape(n, 1 if aosym == "s8" else 2)
    out = torch.empty(shape, dtype=torch.float64, device=_device())
    rc = _native.lib().gscf_packed_eri(atm.ctypes.data, len(atm), bas.ctypes.data, len(bas), env.ctypes.data,
                                       len(env), int(bool(cart)), 1 if aosym == "s8" else 0, out.data_ptr(),
                                       out.numel(), _stream_handle())
    _native.check(rc)
    return out


def packed_jk(eri, density, *, with_j=True, with_k=True):
    """J/K from s8 packed ERIs (backends/native/jk.py:522-546): density
    (..., n, n) float64 or complex128 on the GPU -> (J, K) shaped like it;
    zeros for an unselected operator."""
    if eri.dim() != 1:
        raise ValueError("The native packed J/K operator requires s8 ERIs.")
    _validate_selection(with_j, with_k)
    if not (with_j or with_k):
        return torch.zeros_like(density), torch.zeros_like(density)
    n = density.shape[-1]
    if density.dim() < 2 or tuple(density.shape[-2:]) != (n, n):
        raise ValueError("Packed J/K dimensions do not match")
    # both operands go to the engine's GPU before their pointers are taken (a host
    # pointer handed to the kernels would be dereferenced on the device)
    dev = _device()
    flat = density.reshape(-1, n, n).to(dev)
    count = flat.shape[0]
    real = torch.cat((flat.real, flat.imag), 0) if flat.is_complex() else flat
    real = real.to(torch.float64).contiguous()
    eri = eri.to(dev, torch.float64).contiguous()
    nops = int(bool(with_j)) + int(bool(with_k))
    out = torch.empty((nops,) + tuple(real.shape), dtype=torch.float64, device=real.device)
    _native.check(_native.lib().gscf_packed_jk(eri.data_ptr(), eri.numel(), 1, real.data_ptr(), real.shape[0], n,
                                               out.data_ptr(), int(bool(with_j)), int(bool(with_k)),
                                               _stream_handle()))
    if flat.is_complex():
        out = out[:, :count] + 1j * out[:, count:]
    out = out.reshape(nops, *density.shape)
    return (out[0] if with_j else torch.zeros_like(density),
            out[-1] if with_k else torch.zeros_like(density))


@dataclass(frozen=True)
class IntegralPlan:
    """Plan over a basis topology (the native backend only)."""

    topology: BasisTopology
    backend: str = "native"

    def pack(self, parameters: BasisParameters, *, origin=None):
        return pack(self.topology, parameters, origin=origin)

    def get_jk(self, parameters: BasisParameters, density, *, screening_threshold=0.0, workers=1,
               with_j=True, with_k=True, omega=0.0):
        """J/K of density (..., n, n) with the reference's argument checks.

        When exponents, coefficients, centres or nuclear coordinates of
        `parameters` are tensors that require gradients, the environment is
        packed with differentiable operations (packing.pack_torch), so J/K
        gradients reach them (unscreened plans)."""
        if omega:
            raise NotImplementedError("Range-separated (omega != 0) J/K is not supported by the MI355X engine.")
        if self.backend != "native":
            raise NotImplementedError("get_jk requires the native backend.")
        if tuple(density.shape[-2:]) != (self.topology.nao(),) * 2:
            raise ValueError("Density AO dimensions do not match the plan.")
        leaves = [*parameters.exponents, *parameters.coefficients, parameters.centers, parameters.nuclear_coords]
        if any(isinstance(x, torch.Tensor) and x.requires_grad for x in leaves) and torch.is_grad_enabled():
            atm, bas, env_t = pack_torch(self.topology, parameters)
            return direct_jk(atm, bas, env_t, density, cart=self.topology.cart, cutoff=screening_threshold,
                             workers=workers, with_j=with_j, with_k=with_k)
        params = BasisParameters(tuple(_host(a) for a in parameters.exponents),
                                 tuple(_host(c) for c in parameters.coefficients),
                                 _host(parameters.centers), _host(parameters.nuclear_coords))
        atm, bas, env = self.pack(params)
        coordinates = np.concatenate([np.asarray(params.nuclear_coords).reshape(-1, 3),
                                      np.asarray(params.centers).reshape(-1, 3)])
        return direct_jk(atm, bas, env, density, cart=self.topology.cart, cutoff=screening_threshold,
                         workers=workers, with_j=with_j, with_k=with_k, coordinates=coordinates)


def _host(x):
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def make_plan(topology: BasisTopology, backend: str = "native") -> IntegralPlan:
    if backend != "native":
        raise ValueError("Only the native (MI355X) backend is available.")
    return IntegralPlan(topology, backend)


@dataclass
class NativeDirectBasis:
    plan: IntegralPlan
    parameters: BasisParameters

    @property
    def nao(self) -> int:
        return self.plan.topology.nao()

    def get_jk(self, density, *, screening_threshold=0.0, workers=1, with_j=True, with_k=True, omega=0.0):
        return self.plan.get_jk(self.parameters, density, screening_threshold=screening_threshold,
                                workers=workers, with_j=with_j, with_k=with_k, omega=omega)


def native_jk_workspace_bytes(provider, workers, cutoff, nb: int = 1):
    """Device-memory forecast of a direct J/K plan plus one build of nb
    densities, before any GPU work (mirrors the reference's
    native_jk_workspace_bytes, src/gradscf/integrals/backends/native/jk.py:14-68:
    same arguments, None for a provider without a complete shell topology).
    The bound counts every canonical pair quartet as retained; `workers` is
    validated and otherwise unused (one GPU does the work)."""
    if isinstance(workers, bool) or not isinstance(workers, Integral) or workers < 1:
        raise ValueError("workers must be a positive integer.")
    plan = getattr(provider, "plan", None)
    params = getattr(provider, "parameters", None)
    topology = getattr(plan, "topology", None)
    if topology is None or params is None or not hasattr(topology, "nao"):
        return None
    atm, bas, env = pack(topology, params)
    atm = np.ascontiguousarray(atm, dtype=np.int32)
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    out = ctypes.c_int64()
    _native.check(_native.lib().gscf_workspace_bytes(
        atm.ctypes.data, len(atm), bas.ctypes.data, len(bas), env.ctypes.data, len(env),
        int(bool(topology.cart)), ctypes.c_double(float(cutoff)), int(nb), ctypes.byref(out)))
    return int(out.value)
