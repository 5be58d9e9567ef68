"""Oracle for the direct-SCF J/K path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import
this package, and only as the checker.  The product (gradscf_amd) never imports
it.  It wraps

  * liboracle.so          the project's plain-C restatement of the reference
                          direct-J/K path (boys_rys.c, eri.c, direct_jk.c);
  * _ref/libcint_ref.so   the reference's vendored libcint 6.1.3, compiled from
                          its own sources by oracle/Makefile (when it was built).

Parity status per function is stated in oracle/oracle.h (what pins it).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libcint_ref.so"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None
_ref_loaded = False


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and _ref/libcint_ref.so when ref and /root/reference exist)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if ref and Path("/root/reference").is_dir():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.is_file():
            build()
        _lib = ctypes.CDLL(str(LIB))
        _lib.oracle_last_error.restype = ctypes.c_char_p
    return _lib


def have_ref() -> bool:
    return REF_LIB.is_file()


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())


def select_backend(name: str) -> None:
    """'oracle' -> own Rys ERIs; 'libcint' -> the reference's libcint from _ref."""
    global _ref_loaded
    L = lib()
    if name == "libcint":
        if not _ref_loaded:
            if not have_ref():
                raise RuntimeError("oracle/_ref/libcint_ref.so is not built")
            _check(L.oracle_use_libcint(str(REF_LIB).encode()))
            _ref_loaded = True
        L.oracle_select_backend(1)
    elif name == "oracle":
        L.oracle_select_backend(0)
    else:
        raise ValueError(name)


class _Backend:
    def __init__(self, name):
        self.name = name

    def __enter__(self):
        select_backend(self.name)

    def __exit__(self, *exc):
        select_backend("oracle")


def backend(name: str) -> _Backend:
    return _Backend(name)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def boys(mmax: int, x: float) -> np.ndarray:
    f = np.zeros(mmax + 1)
    lib().oracle_boys(ctypes.c_int(mmax), ctypes.c_double(x), _p(f))
    return f


def rys_roots(n: int, x: float):
    t2, w = np.zeros(n), np.zeros(n)
    _check(lib().oracle_rys_roots(ctypes.c_int(n), ctypes.c_double(x), _p(t2), _p(w)))
    return t2, w


def nao(bas, cart: bool) -> int:
    bas = _i32(bas)
    return int(lib().oracle_nao(_p(bas, _ip), len(bas), int(cart)))


def ao_offsets(bas, cart):
    ao = [0]
    for row in np.asarray(bas):
        l, nc = int(row[1]), int(row[3])
        ao.append(ao[-1] + nc * ((l + 1) * (l + 2) // 2 if cart else 2 * l + 1))
    return np.array(ao)


def eri_block(shls, atm, bas, env, cart=True, deriv=0) -> np.ndarray:
    """Contracted block in libcint layout, returned as array [d0, d1, d2, d3] (C order)
    with an extra leading axis of 3 for deriv=1 (electron-coordinate nabla on index 0)."""
    atm, bas, env = _i32(atm), _i32(bas), _f64(env)
    ao = ao_offsets(bas, cart)
    d = [int(ao[s + 1] - ao[s]) for s in shls]
    ncomp = 3 if deriv else 1
    out = np.zeros(ncomp * d[0] * d[1] * d[2] * d[3])
    sh = _i32(shls)
    _check(lib().oracle_eri_block(_p(sh, _ip), _p(atm, _ip), len(atm), _p(bas, _ip), len(bas),
                                  _p(env), int(cart), int(deriv), _p(out)))
    # libcint layout: first index fastest -> reshape reversed then transpose
    blk = out.reshape(ncomp, d[3], d[2], d[1], d[0]).transpose(0, 4, 3, 2, 1)
    return blk if deriv else blk[0]


def full_eri(atm, bas, env, cart=True) -> np.ndarray:
    """Full s1 (pq|rs) tensor assembled from oracle blocks (tiny systems only)."""
    ao = ao_offsets(bas, cart)
    n = int(ao[-1])
    eri = np.zeros((n, n, n, n))
    nb = len(bas)
    for i in range(nb):
        for j in range(nb):
            for k in range(nb):
                for l in range(nb):
                    eri[ao[i]:ao[i + 1], ao[j]:ao[j + 1], ao[k]:ao[k + 1], ao[l]:ao[l + 1]] = \
                        eri_block((i, j, k, l), atm, bas, env, cart)
    return eri


def schwarz(atm, bas, env, cart=True) -> np.ndarray:
    atm, bas, env = _i32(atm), _i32(bas), _f64(env)
    q = np.zeros(len(bas) * len(bas))
    _check(lib().oracle_schwarz(_p(atm, _ip), len(atm), _p(bas, _ip), len(bas), _p(env),
                                int(cart), _p(q)))
    return q.reshape(len(bas), len(bas))


def direct_jk(atm, bas, env, dm, *, cart=True, cutoff=0.0, workers=1, with_j=True,
              with_k=True) -> np.ndarray:
    """Returns out[nops, nb, n, n] exactly like the reference FFI target."""
    atm, bas, env = _i32(atm), _i32(bas), _f64(env)
    dm = _f64(dm)
    squeeze = dm.ndim == 2
    if squeeze:
        dm = dm[None]
    nb, n = dm.shape[0], dm.shape[1]
    nops = int(with_j) + int(with_k)
    out = np.zeros((nops, nb, n, n))
    _check(lib().oracle_direct_jk(_p(atm, _ip), len(atm), _p(bas, _ip), len(bas), _p(env),
                                  len(env), _p(dm), nb, _p(out), int(cart), ctypes.c_double(cutoff),
                                  int(workers), int(with_j), int(with_k)))
    return out


def int1e(atm, bas, env, cart=True):
    atm, bas, env = _i32(atm), _i32(bas), _f64(env)
    n = nao(bas, cart)
    s, t, v = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    _check(lib().oracle_int1e(_p(atm, _ip), len(atm), _p(bas, _ip), len(bas), _p(env), int(cart),
                              _p(s), _p(t), _p(v)))
    return s, t, v


def jk_geometry_vjp(atm, bas, env, dm, cot_j=None, cot_k=None, *, cart=True) -> np.ndarray:
    """grad[natm, 3] of sum(cot_j*J) + sum(cot_k*K) w.r.t. the atm-row coordinates."""
    atm, bas, env, dm = _i32(atm), _i32(bas), _f64(env), _f64(dm)
    g = np.zeros(3 * len(atm))
    cj = None if cot_j is None else _f64(cot_j)
    ck = None if cot_k is None else _f64(cot_k)
    _check(lib().oracle_jk_geometry_vjp(_p(atm, _ip), len(atm), _p(bas, _ip), len(bas), _p(env),
                                        len(env), _p(dm), None if cj is None else _p(cj),
                                        None if ck is None else _p(ck), int(cart), _p(g)))
    return g.reshape(len(atm), 3)


def libcint_block(shls, atm, bas, env, cart=True, intor="int2e"):
    """The reference libcint's own block (needs _ref), same array convention as eri_block."""
    select_backend("libcint")
    select_backend("oracle")
    h = ctypes.CDLL(str(REF_LIB))
    atm, bas, env = _i32(atm), _i32(bas), _f64(env)
    ao = ao_offsets(bas, cart)
    d = [int(ao[s + 1] - ao[s]) for s in shls]
    ncomp = 3 if intor == "int2e_ip1" else 1
    out = np.zeros(ncomp * int(np.prod(d)))
    f = getattr(h, f"{intor}_{'cart' if cart else 'sph'}")
    sh = _i32(shls)
    f(_p(out), None, _p(sh, _ip), _p(atm, _ip), len(atm), _p(bas, _ip), len(bas), _p(env), None, None)
    blk = out.reshape(ncomp, d[3], d[2], d[1], d[0]).transpose(0, 4, 3, 2, 1)
    return blk if ncomp == 3 else blk[0]


def libcint_rys_roots(n: int, x: float):
    h = ctypes.CDLL(str(REF_LIB))
    u, w = np.zeros(n), np.zeros(n)
    h.CINTrys_roots(ctypes.c_int(n), ctypes.c_double(x), _p(u), _p(w))
    return u, w


def libcint_boys(mmax: int, x: float) -> np.ndarray:
    h = ctypes.CDLL(str(REF_LIB))
    f = np.zeros(mmax + 1)
    h.gamma_inc_like(_p(f), ctypes.c_double(x), ctypes.c_int(mmax))
    return f


def libcint_int1e(atm, bas, env, cart=True):
    h = ctypes.CDLL(str(REF_LIB))
    atm, bas, env = _i32(atm), _i32(bas), _f64(env)
    ao = ao_offsets(bas, cart)
    n = int(ao[-1])
    mats = []
    for name in ("int1e_ovlp", "int1e_kin", "int1e_nuc"):
        f = getattr(h, f"{name}_{'cart' if cart else 'sph'}")
        m = np.zeros((n, n))
        for i in range(len(bas)):
            for j in range(len(bas)):
                di, dj = ao[i + 1] - ao[i], ao[j + 1] - ao[j]
                out = np.zeros(di * dj)
                sh = _i32([i, j])
                f(_p(out), None, _p(sh, _ip), _p(atm, _ip), len(atm), _p(bas, _ip), len(bas), _p(env),
                  None, None)
                m[ao[i]:ao[i + 1], ao[j]:ao[j + 1]] = out.reshape(dj, di).T
        mats.append(m)
    return tuple(mats)


def libcint_aux_block(shls, atm, bas, env, cart=True):
    """The reference libcint's int2c2e (two shells) or int3c2e (three shells)
    block as an array [d0, d1(, d2)] (libcint's first-index-fastest layout
    transposed to C order); needs _ref."""
    h = ctypes.CDLL(str(REF_LIB))
    atm, bas, env = _i32(atm), _i32(bas), _f64(env)
    ao = ao_offsets(bas, cart)
    d = [int(ao[s + 1] - ao[s]) for s in shls]
    out = np.zeros(int(np.prod(d)))
    f = getattr(h, f"{'int3c2e' if len(shls) == 3 else 'int2c2e'}_{'cart' if cart else 'sph'}")
    sh = _i32(shls)
    f(_p(out), None, _p(sh, _ip), _p(atm, _ip), len(atm), _p(bas, _ip), len(bas), _p(env), None, None)
    return out.reshape(d[::-1]).transpose(tuple(range(len(d)))[::-1])


def density_fitting(atm, bas, env, *, layout, cart, split):
    """Auxiliary integrals in the reference's compact layouts, restating the
    loops of GradSCF's density_fitting (csrc/density_fitting.cc:5-34) over the
    reference libcint's int2c2e / int3c2e blocks (needs _ref):
      layout 2: (P|Q), m x m, every shell pair i >= j written to both halves;
      layout 3: (ab|P), m x n(n+1)/2 at pair_index(a, b) = a(a+1)/2 + b, a >= b,
    with orbital shells [0, split) and auxiliary shells [split, nbas)."""
    bas = _i32(bas)
    nbas = len(bas)
    if split < 1 or split >= nbas:
        raise ValueError("Invalid orbital/auxiliary shell split")
    ao = ao_offsets(bas, cart)
    n = int(ao[split])
    m = int(ao[-1]) - n
    npair = n * (n + 1) // 2
    if layout == 2:
        out = np.zeros((m, m))
        for i in range(split, nbas):
            for j in range(split, i + 1):
                blk = libcint_aux_block((i, j), atm, bas, env, cart)
                for p in range(blk.shape[0]):
                    for q in range(blk.shape[1]):
                        a, b = ao[i] + p - n, ao[j] + q - n
                        out[a, b] = out[b, a] = blk[p, q]
        return out
    out = np.zeros((m, npair))
    for i in range(split):
        for j in range(i + 1):
            for k in range(split, nbas):
                blk = libcint_aux_block((i, j, k), atm, bas, env, cart)
                for p in range(blk.shape[0]):
                    for q in range(blk.shape[1]):
                        a, b = ao[i] + p, ao[j] + q
                        if a < b:
                            continue
                        out[ao[k] - n:ao[k + 1] - n, a * (a + 1) // 2 + b] = blk[p, q, :]
    return out
