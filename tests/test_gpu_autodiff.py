"""Differentiation of direct J/K and the Python call surface, on the GPU.

Mirrors the reference's tests of its direct-J/K primitive
(tests/integrals/test_native_direct_parallel.py): the same two-centre H2
basis with a general contraction, a p and a d shell ("H 0 0 0; H 0.1 0.2
1.1", spherical and Cartesian), nonsymmetric batched densities, J/K checked
against contractions of the full ERI tensor (the oracle's, as the reference
uses its own full tensor), and

  * :105-129  density JVP, VJP of 0.7 sum(J^2) + 0.3 sum(K^2) and its
              Hessian-vector product;
  * :132-165  operator selection keeps nonsymmetric densities and AD;
plus the coordinate VJP against the oracle's geometry VJP and central
differences, exponent/coefficient gradients through IntegralPlan.get_jk
against central differences, IntegralPlan.get_jk / NativeDirectBasis.get_jk
(with explicit coordinates) against the oracle, and the one-shot C-ABI
gscf_direct_jk on device buffers against the oracle.
Tolerances: 2e-10 abs + 1e-10 rel on J/K (the reference's), 2e-9 on
gradients of the quadratic loss.
"""
import ctypes

import numpy as np
import pytest

from tests import systems

pytestmark = pytest.mark.gpu

SPEC = [[0, [2.0, 0.8, -0.2], [0.4, 0.3, 1.0]], [1, [0.8, 1.0]], [2, [0.6, 1.0]]]
ATOMS = "H 0 0 0; H 0.1 0.2 1.1"


@pytest.fixture(scope="module", params=[False, True], ids=["spherical", "cartesian"])
def basis(request, oracle_mod):
    from gradscf_amd.basis import prepare_basis
    from gradscf_amd.packing import pack
    topo, par = prepare_basis(ATOMS, SPEC, cart=request.param)
    atm, bas, env = pack(topo, par)
    eri = oracle_mod.full_eri(atm, bas, env, cart=request.param)
    return topo, par, atm, bas, env, eri


def _cuda(x):
    import torch
    return torch.as_tensor(np.asarray(x), dtype=torch.float64, device="cuda")


def _action(basis, d, **kw):
    from gradscf_amd.jk import direct_jk
    topo, _, atm, bas, env, _ = basis
    return direct_jk(atm, bas, env, d, cart=topo.cart, **kw)


def _full(basis, d):
    import torch
    eri = torch.as_tensor(basis[-1], dtype=torch.float64, device=d.device).to(d.dtype)
    return torch.einsum("pqrs,...rs->...pq", eri, d), torch.einsum("prqs,...rs->...pq", eri, d)


def _close(a, b, atol=2e-10, rtol=1e-10):
    a, b = a.detach().cpu().numpy(), b.detach().cpu().numpy()
    np.testing.assert_allclose(a, b, atol=atol, rtol=rtol)


def test_full_tensor_batched_nonsymmetric(basis):
    n = basis[0].nao()
    d = _cuda(np.random.default_rng(801).normal(size=(4, n, n)))
    for got, ref in zip(_action(basis, d), _full(basis, d)):
        _close(got, ref)


def test_density_jvp_vjp_and_hvp(basis):
    import torch
    import torch.autograd.forward_ad as fwad
    n = basis[0].nao()
    rng = np.random.default_rng(804)
    d, t = (_cuda(x) for x in rng.normal(size=(2, 4, n, n)))
    with fwad.dual_level():
        out = _action(basis, fwad.make_dual(d, t))
        tangents = [fwad.unpack_dual(o).tangent for o in out]
    for got, ref in zip(tangents, _full(basis, t)):
        _close(got, ref)

    def loss(action, x):
        j, k = action(x)
        return 0.7 * torch.sum(j ** 2) + 0.3 * torch.sum(k ** 2)

    def grad_of(action, x, create_graph=False):
        x = x.detach().requires_grad_(True)
        return x, torch.autograd.grad(loss(action, x), x, create_graph=create_graph)[0]

    full = lambda x: _full(basis, x)
    act = lambda x: _action(basis, x)
    _, g = grad_of(act, d)
    _, g_ref = grad_of(full, d)
    _close(g, g_ref, atol=2e-9, rtol=2e-10)
    x, g = grad_of(act, d, create_graph=True)
    hvp = torch.autograd.grad(g, x, grad_outputs=t)[0]
    x, g = grad_of(full, d, create_graph=True)
    hvp_ref = torch.autograd.grad(g, x, grad_outputs=t)[0]
    _close(hvp, hvp_ref, atol=2e-9, rtol=2e-10)


@pytest.mark.parametrize("with_j,with_k", [(True, True), (True, False), (False, True), (False, False)])
def test_selected_operators_keep_nonsymmetric_density_and_ad(basis, with_j, with_k):
    import torch
    n = basis[0].nao()
    rng = np.random.default_rng(806)
    d0, cot = (_cuda(x) for x in rng.normal(size=(2, 4, n, n)))
    d = d0.clone().requires_grad_(True)
    j, k = _action(basis, d, with_j=with_j, with_k=with_k)
    jr, kr = _full(basis, d0)
    _close(j, jr if with_j else torch.zeros_like(jr))
    _close(k, kr if with_k else torch.zeros_like(kr))
    if not (with_j or with_k):
        assert not j.requires_grad and not k.requires_grad
        return
    (torch.sum(cot * j) + torch.sum(cot * k)).backward()
    dr = d0.clone().requires_grad_(True)
    jr, kr = _full(basis, dr)
    (torch.sum(cot * jr) * with_j + torch.sum(cot * kr) * with_k).backward()
    _close(d.grad, dr.grad, atol=2e-9, rtol=2e-10)


def test_complex_density_gradient(basis):
    """A complex density goes through the real operator on its real and
    imaginary parts; the gradient of a real loss follows torch's convention."""
    import torch
    n = basis[0].nao()
    rng = np.random.default_rng(807)
    z0 = torch.complex(*(_cuda(x) for x in rng.normal(size=(2, 2, n, n))))
    z = z0.clone().requires_grad_(True)
    j, k = _action(basis, z)
    (torch.sum(j.abs() ** 2) + 0.5 * torch.sum(k.real * k.imag)).backward()
    zr = z0.clone().requires_grad_(True)
    jr, kr = _full(basis, zr)
    (torch.sum(jr.abs() ** 2) + 0.5 * torch.sum(kr.real * kr.imag)).backward()
    _close(z.grad.real, zr.grad.real, atol=2e-9, rtol=2e-10)
    _close(z.grad.imag, zr.grad.imag, atol=2e-9, rtol=2e-10)


def test_coordinate_gradient_matches_oracle_and_fd(basis, oracle_mod):
    import torch
    from gradscf_amd.jk import direct_jk
    topo, par, atm, bas, env, _ = basis
    n = topo.nao()
    rng = np.random.default_rng(808)
    d = rng.normal(size=(n, n))
    cj, ck = rng.normal(size=(2, n, n))
    coords0 = np.concatenate([par.nuclear_coords, par.centers]).astype(np.float64)
    coords = torch.tensor(coords0, requires_grad=True)
    j, k = direct_jk(atm, bas, env, _cuda(d)[None], cart=topo.cart, coordinates=coords)
    (torch.sum(_cuda(cj) * j[0]) + torch.sum(_cuda(ck) * k[0])).backward()
    ref = oracle_mod.jk_geometry_vjp(atm, bas, env, d, cj, ck, cart=topo.cart)
    got = coords.grad.numpy()
    assert np.max(np.abs(got - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))

    def f(c):
        j, k = direct_jk(atm, bas, env, _cuda(d), cart=topo.cart, coordinates=c)
        return float(torch.sum(_cuda(cj) * j) + torch.sum(_cuda(ck) * k))

    h = 1e-4
    for row, ax in ((0, 2), (2, 0), (len(atm) - 1, 1)):
        cp, cm = coords0.copy(), coords0.copy()
        cp[row, ax] += h
        cm[row, ax] -= h
        fd = (f(cp) - f(cm)) / (2 * h)
        assert abs(fd - got[row, ax]) < 1e-6 * max(1.0, abs(fd)), (row, ax, fd, got[row, ax])
    # screened operators refuse geometry AD, as the reference does
    coords = torch.tensor(coords0, requires_grad=True)
    j, _ = direct_jk(atm, bas, env, _cuda(d), cart=topo.cart, coordinates=coords, cutoff=1e-12)
    with pytest.raises(NotImplementedError, match="cutoff=0"):
        torch.sum(j).backward()


def test_basis_parameter_gradients_through_get_jk(basis):
    """Raw exponents, raw coefficients (through the normalisation), shell
    centres and nuclear coordinates as leaf tensors of IntegralPlan.get_jk:
    gradients of sum(cJ J) + sum(cK K) against central differences."""
    import torch
    from gradscf_amd.basis import BasisParameters
    from gradscf_amd.jk import make_plan
    topo, par = basis[0], basis[1]
    plan = make_plan(topo)
    n = topo.nao()
    rng = np.random.default_rng(809)
    d = _cuda(rng.normal(size=(n, n)))
    cj, ck = (_cuda(x) for x in rng.normal(size=(2, n, n)))
    exps = tuple(torch.tensor(np.asarray(a, dtype=float), requires_grad=True) for a in par.exponents)
    coefs = tuple(torch.tensor(np.asarray(c, dtype=float), requires_grad=True) for c in par.coefficients)
    cent = torch.tensor(np.asarray(par.centers, dtype=float), requires_grad=True)
    nuc = torch.tensor(np.asarray(par.nuclear_coords, dtype=float), requires_grad=True)
    j, k = plan.get_jk(BasisParameters(exps, coefs, cent, nuc), d)
    (torch.sum(cj * j) + torch.sum(ck * k)).backward()

    def f(p):
        j, k = plan.get_jk(p, d)
        return float(torch.sum(cj * j) + torch.sum(ck * k))

    def moved(kind, s, idx, h):
        ex = [np.array(a, dtype=float) for a in par.exponents]
        co = [np.array(c, dtype=float) for c in par.coefficients]
        ce = np.array(par.centers, dtype=float)
        if kind == "exp":
            ex[s][idx] += h
        elif kind == "coef":
            co[s][idx] += h
        else:
            ce[s, idx] += h
        return BasisParameters(tuple(ex), tuple(co), ce, np.asarray(par.nuclear_coords))

    checks = [("exp", 0, 0, exps[0].grad[0]), ("exp", 2, 0, exps[2].grad[0]),
              ("coef", 0, (1, 1), coefs[0].grad[1, 1]), ("coef", 0, (0, 0), coefs[0].grad[0, 0]),
              ("cent", 2, 2, cent.grad[2, 2]), ("cent", 4, 0, cent.grad[4, 0])]
    for kind, s, idx, an in checks:
        h = 1e-5 * (par.exponents[s][idx] if kind == "exp" else 1.0)
        fd = (f(moved(kind, s, idx, h)) - f(moved(kind, s, idx, -h))) / (2 * h)
        assert abs(fd - float(an)) < 1e-6 * max(1.0, abs(fd)), (kind, s, idx, fd, float(an))
    # basis centres carry every derivative here (the nuclei hold no basis functions)
    assert float(nuc.grad.abs().max()) == 0.0


@pytest.mark.parametrize("cart", [True, False])
def test_integral_plan_get_jk_matches_oracle(oracle_mod, cart):
    """IntegralPlan.get_jk and NativeDirectBasis.get_jk (plan.py:71-92,
    jk.py:98-107): coordinates passed separately from the packed tables."""
    import torch
    from gradscf_amd.basis import prepare_basis
    from gradscf_amd.jk import NativeDirectBasis, make_plan
    from gradscf_amd.packing import pack
    topo, par = prepare_basis(systems.water_dimer(), "cc-pvdz", cart=cart)
    atm, bas, env = pack(topo, par)
    plan = make_plan(topo)
    d = systems.random_density(topo.nao(), nb=2, seed=810, symmetric=False)
    ref = oracle_mod.direct_jk(atm, bas, env, d, cart=cart, cutoff=1e-12, workers=4)
    j, k = plan.get_jk(par, _cuda(d), screening_threshold=1e-12)
    assert np.max(np.abs(j.cpu().numpy() - ref[0])) < 1e-10
    assert np.max(np.abs(k.cpu().numpy() - ref[1])) < 1e-10
    nb = NativeDirectBasis(plan, par)
    assert nb.nao == topo.nao()
    j2, k2 = nb.get_jk(_cuda(d[0]), screening_threshold=1e-12, with_j=False)
    assert torch.count_nonzero(j2) == 0
    assert np.max(np.abs(k2.cpu().numpy() - ref[1, 0])) < 1e-10
    with pytest.raises(NotImplementedError):
        plan.get_jk(par, _cuda(d), omega=0.3)
    with pytest.raises(ValueError, match="AO dimensions"):
        plan.get_jk(par, _cuda(d[:, :-1, :-1]))


@pytest.mark.parametrize("cart,with_j,with_k", [(True, 1, 1), (False, 1, 1), (True, 1, 0), (False, 0, 1)])
def test_one_shot_c_abi_on_device_buffers(oracle_mod, cart, with_j, with_k):
    """gscf_direct_jk with the argument list of GradSCFDirectJK
    (direct_jk.cc:97-102), called through ctypes with device buffers: two
    nonsymmetric densities, output [op][b][n][n]."""
    import torch
    from gradscf_amd import _native
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=cart)
    atm = np.ascontiguousarray(atm, dtype=np.int32)
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    n = oracle_mod.nao(bas, cart)
    d = systems.random_density(n, nb=2, seed=811, symmetric=False)
    ref = oracle_mod.direct_jk(atm, bas, env, d, cart=cart, cutoff=1e-12, workers=4,
                               with_j=bool(with_j), with_k=bool(with_k))
    nops = with_j + with_k
    dd = _cuda(d).contiguous()
    out = torch.full((nops, 2, n, n), float("nan"), dtype=torch.float64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = _native.lib().gscf_direct_jk(atm.ctypes.data, len(atm), bas.ctypes.data, len(bas), env.ctypes.data,
                                      len(env), dd.data_ptr(), 2, n, out.data_ptr(), int(cart),
                                      ctypes.c_double(1e-12), 4, with_j, with_k, stream)
    _native.check(rc)
    torch.cuda.synchronize()
    assert np.max(np.abs(out.cpu().numpy() - ref)) < 1e-10
    # a shape the tables do not give is reported as Internal, like the reference
    rc = _native.lib().gscf_direct_jk(atm.ctypes.data, len(atm), bas.ctypes.data, len(bas), env.ctypes.data,
                                      len(env), dd.data_ptr(), 2, n - 1, out.data_ptr(), int(cart),
                                      ctypes.c_double(1e-12), 4, with_j, with_k, stream)
    assert rc == _native.GSCF_INTERNAL
