"""Packed ERIs and packed J/K (the reference's compact-integral and packed J/K
targets, csrc/packed_eri.cc + csrc/packed_jk.cc): the packed values against
the oracle's full ERI tensor, packed J/K against the reference test's own
construction (every s1 element looked up in the s8 triangle, then contracted,
tests/integrals/test_native_packed_selection.py:11-32) at its 3e-12
tolerance, and packed J/K of packed ERIs against direct J/K at the 1e-10 bar.
"""
import numpy as np
import pytest

from gradscf_amd import molecule as M
from tests import systems

pytestmark = pytest.mark.gpu


def _s8_indices(n):
    ids = np.arange(n)
    hi, lo = np.maximum(ids[:, None], ids), np.minimum(ids[:, None], ids)
    pairs = hi * (hi + 1) // 2 + lo
    hi = np.maximum(pairs[:, :, None, None], pairs[None, None, :, :])
    lo = np.minimum(pairs[:, :, None, None], pairs[None, None, :, :])
    return hi * (hi + 1) // 2 + lo


@pytest.mark.parametrize("spec,basis,cart", [("H 0 0 0; H 0 0 1.4", "6-31g", True),
                                             ("water", "6-31g", False), ("water", "6-31g*", True)])
def test_packed_eri_matches_oracle_full_tensor(oracle_mod, spec, basis, cart):
    from gradscf_amd.jk import packed_eri, packed_eri_shape
    mol = systems.water() if spec == "water" else M.parse_molecule_spec(spec)
    _, _, atm, bas, env = systems.tables(mol, basis, cart=cart)
    full = oracle_mod.full_eri(atm, bas, env, cart=cart)
    n = full.shape[0]
    s8 = packed_eri(atm, bas, env, cart=cart, aosym="s8").cpu().numpy()
    assert s8.shape == packed_eri_shape(n, 1)
    assert np.max(np.abs(s8[_s8_indices(n)] - full)) < 1e-12
    s4 = packed_eri(atm, bas, env, cart=cart, aosym="s4").cpu().numpy()
    assert s4.shape == packed_eri_shape(n, 2)
    rows, cols = np.tril_indices(n)
    assert np.max(np.abs(s4 - full[rows, cols][:, rows, cols])) < 1e-12


@pytest.mark.parametrize("with_j,with_k", [(True, True), (True, False), (False, True), (False, False)])
def test_packed_jk_selected_operators_match_full_tensor_contraction(with_j, with_k):
    import torch
    from gradscf_amd.jk import packed_eri, packed_jk
    topo, _, atm, bas, env = systems.tables(M.parse_molecule_spec("H 0 0 0; H 0 0 1.4"), "6-31g", cart=True)
    eri = packed_eri(atm, bas, env, cart=True)
    n = topo.nao()
    full = eri.cpu().numpy()[_s8_indices(n)]
    density = np.random.default_rng(920).normal(size=(2, n, n))
    j, k = packed_jk(eri, torch.as_tensor(density, device="cuda"), with_j=with_j, with_k=with_k)
    ref_j = np.einsum("pqrs,brs->bpq", full, density) if with_j else np.zeros_like(density)
    ref_k = np.einsum("prqs,brs->bpq", full, density) if with_k else np.zeros_like(density)
    assert np.max(np.abs(j.cpu().numpy() - ref_j)) < 3e-12
    assert np.max(np.abs(k.cpu().numpy() - ref_k)) < 3e-12


def test_packed_jk_of_packed_eri_equals_direct_jk():
    import torch
    from gradscf_amd.jk import direct_jk, packed_eri, packed_jk
    topo, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=False)
    eri = packed_eri(atm, bas, env, cart=False)
    n = topo.nao()
    d = torch.as_tensor(systems.random_density(n, nb=2, seed=9, symmetric=False), device="cuda")
    jp, kp = packed_jk(eri, d)
    jd, kd = direct_jk(atm, bas, env, d, cart=False, cutoff=0.0)
    assert float((jp - jd).abs().max()) < 1e-10
    assert float((kp - kd).abs().max()) < 1e-10


def test_packed_entry_points_reject_bad_shapes_as_internal():
    import torch
    from gradscf_amd import _native
    from gradscf_amd.jk import packed_eri, packed_jk
    _, _, atm, bas, env = systems.tables(systems.water(), "6-31g", cart=True)
    eri = packed_eri(atm, bas, env, cart=True)
    d = torch.zeros((1, 13, 13), dtype=torch.float64, device="cuda")
    out = torch.zeros((2, 1, 13, 13), dtype=torch.float64, device="cuda")
    L = _native.lib()
    rc = L.gscf_packed_jk(eri.data_ptr(), eri.numel() - 1, 1, d.data_ptr(), 1, 13, out.data_ptr(), 1, 1, None)
    assert rc == _native.GSCF_INTERNAL
    assert L.gscf_last_error().decode() == "Packed J/K dimensions do not match"
    rc = L.gscf_packed_jk(eri.data_ptr(), eri.numel(), 1, d.data_ptr(), 1, 13, out.data_ptr(), 0, 0, None)
    assert rc == _native.GSCF_INTERNAL
    assert L.gscf_last_error().decode() == "Packed J/K requires at least one selected operator"
    a, b = (np.ascontiguousarray(x, dtype=np.int32) for x in (atm, bas))
    e = np.ascontiguousarray(env, dtype=np.float64)
    rc = L.gscf_packed_eri(a.ctypes.data, len(a), b.ctypes.data, len(b), e.ctypes.data, len(e), 1, 1,
                           eri.data_ptr(), eri.numel() + 1, None)
    assert rc == _native.GSCF_INTERNAL
    assert L.gscf_last_error().decode() == "Packed ERI shape mismatch"
    with pytest.raises(ValueError):
        packed_jk(eri.reshape(1, -1), d)


def test_packed_jk_complex_density_is_linear():
    """Complex densities are contracted as their real and imaginary parts
    (backends/native/jk.py:530-546)."""
    import torch
    from gradscf_amd.jk import packed_eri, packed_jk
    topo, _, atm, bas, env = systems.tables(systems.water(), "6-31g", cart=True)
    eri = packed_eri(atm, bas, env, cart=True)
    n = topo.nao()
    re = torch.as_tensor(systems.random_density(n, nb=2, seed=11, symmetric=False), device="cuda")
    im = torch.as_tensor(systems.random_density(n, nb=2, seed=12, symmetric=False), device="cuda")
    j, k = packed_jk(eri, torch.complex(re, im))
    jr, kr = packed_jk(eri, re)
    ji, ki = packed_jk(eri, im)
    assert j.is_complex() and j.shape == re.shape
    assert float((j - torch.complex(jr, ji)).abs().max()) < 1e-12
    assert float((k - torch.complex(kr, ki)).abs().max()) < 1e-12
