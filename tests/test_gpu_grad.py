"""Nuclear gradients on the GPU: the one-electron gradient kernel against
central finite differences of the GPU one-electron integrals (themselves pinned
to the oracle), and the full RHF gradient against central finite differences
of converged SCF energies (step 1e-4 bohr; the target bound is 1e-6 Eh/bohr),
plus the translational sum rule.  Geometry: the reference's published water
(O 0 0 0; H 0 +-1.43 1.11 bohr)."""
import numpy as np
import pytest

from gradscf_amd import molecule as M
from tests import systems

pytestmark = pytest.mark.gpu

WATER = "O 0 0 0; H 0 1.43 1.11; H 0 -1.43 1.11"


def _shift(atm, env, rows, axis, h):
    e = np.array(env, copy=True)
    for r in rows:
        e[int(atm[r, 1]) + axis] += h
    return e


@pytest.mark.parametrize("basis,cart", [("6-31g*", True), ("cc-pvdz", False)])
def test_int1e_gradient_rows_match_fd(basis, cart):
    import torch
    from gradscf_amd.jk import DirectJKPlan
    spec = M.parse_molecule_spec(WATER, unit="Bohr")
    _, _, atm, bas, env = systems.tables(spec, basis, cart=cart)
    plan = DirectJKPlan(atm, bas, env, cart=cart)
    n = plan.nao
    d = torch.as_tensor(systems.random_density(n, seed=81)[0], device="cuda")
    w = torch.as_tensor(systems.random_density(n, seed=82)[0], device="cuda")
    g = plan.int1e_gradient(d, w).cpu().numpy()

    def e1(env_):
        s, t, v = DirectJKPlan(atm, bas, env_, cart=cart).int1e()
        return float((d * (t + v)).sum() - (w * s).sum())

    nnuc = 3
    d_shell = int(np.nonzero(bas[:, 1] == 2)[0][0])
    rows = [0, 1, nnuc, nnuc + d_shell, len(atm) - 1]   # O and H nuclei, O 1s, O d, last H shell
    # a single basis centre moved off its nucleus: the O 1s shell (width ~0.01 bohr)
    # needs a small step; five-point stencil, truncation O(h^4)
    h = 2e-5
    for r in rows:
        for ax in range(3):
            f = [e1(_shift(atm, env, [r], ax, k * h)) for k in (-2, -1, 1, 2)]
            fd = (f[0] - 8 * f[1] + 8 * f[2] - f[3]) / (12 * h)
            assert abs(fd - g[r, ax]) < 1e-7 * max(1.0, abs(fd)), (r, ax, fd, g[r, ax])


def _enuc(z, r):
    z = np.asarray(z, dtype=np.float64)
    return sum(z[i] * z[j] / np.linalg.norm(r[i] - r[j]) for i in range(len(z)) for j in range(i))


def _scf_energy(atm, bas, env, spec, cart):
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.scf import rhf
    res = rhf(DirectJKPlan(atm, bas, env, cart=cart), spec.nelectron, 0.0, conv_tol=1e-12,
              conv_tol_density=1e-10, conv_tol_grad=1e-9)
    assert res.converged
    return res


@pytest.mark.parametrize("basis,cart", [("6-31g*", True), ("cc-pvdz", False)])
def test_rhf_gradient_matches_finite_difference(basis, cart):
    from gradscf_amd.basis import prepare_basis
    from gradscf_amd.grad import nuclear_repulsion_gradient, rhf_gradient, shell_nuclei
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.packing import pack
    spec = M.parse_molecule_spec(WATER, unit="Bohr")
    topo, par = prepare_basis(spec, basis, cart=cart)
    atm, bas, env = pack(topo, par)
    sh = shell_nuclei(par.centers, par.nuclear_coords)
    res = _scf_energy(atm, bas, env, spec, cart)
    plan = DirectJKPlan(atm, bas, env, cart=cart, cutoff=0.0)
    g = rhf_gradient(plan, res.density, res.mo_coeff, res.mo_energy, spec.nelectron // 2,
                     spec.charges, spec.coords_bohr, sh)
    assert np.max(np.abs(g.sum(axis=0))) < 1e-8   # translational invariance
    nnuc = len(spec.charges)
    h = 1e-4
    for a in range(nnuc):
        rows = [a] + [nnuc + i for i in np.nonzero(sh == a)[0]]
        for ax in range(3):
            ep = _scf_energy(atm, bas, _shift(atm, env, rows, ax, h), spec, cart).energy
            em = _scf_energy(atm, bas, _shift(atm, env, rows, ax, -h), spec, cart).energy
            coords_p = spec.coords_bohr.copy()
            coords_p[a, ax] += h
            coords_m = spec.coords_bohr.copy()
            coords_m[a, ax] -= h
            fd = (ep - em) / (2 * h)
            # electronic energies above; nuclear repulsion difference analytically
            fd += (_enuc(spec.charges, coords_p) - _enuc(spec.charges, coords_m)) / (2 * h)
            assert abs(fd - g[a, ax]) < 1e-6, (a, ax, fd, g[a, ax])
    assert np.allclose(nuclear_repulsion_gradient(spec.charges, spec.coords_bohr).sum(axis=0), 0.0, atol=1e-12)


def _slots(bas, shells):
    """(env index, relative step) of every exponent and coefficient slot of the given shells."""
    out = []
    for s in shells:
        row = bas[s]
        for p in range(int(row[2])):
            out.append((int(row[5]) + p, "exp"))
        for k in range(int(row[2]) * int(row[3])):
            out.append((int(row[6]) + k, "coef"))
    return out


def _fd_env(fun, env, idx, kind):
    h = 1e-5 * abs(env[idx]) if kind == "exp" else 1e-5
    ep, em = np.array(env, copy=True), np.array(env, copy=True)
    ep[idx] += h
    em[idx] -= h
    return (fun(ep) - fun(em)) / (2 * h)


@pytest.mark.parametrize("basis,cart", [("6-31g*", True), ("cc-pvdz", False)])
def test_jk_param_vjp_matches_fd(basis, cart):
    """d/d env [sum(cJ*J(D)) + sum(cK*K(D))] at exponent and coefficient slots
    (general contraction included for cc-pVDZ) vs central differences of the
    GPU J/K built from displaced env."""
    import torch
    from gradscf_amd.jk import DirectJKPlan
    spec = M.parse_molecule_spec(WATER, unit="Bohr")
    _, _, atm, bas, env = systems.tables(spec, basis, cart=cart)
    plan = DirectJKPlan(atm, bas, env, cart=cart)
    n = plan.nao
    d = torch.as_tensor(systems.random_density(n, seed=91)[0], device="cuda")
    cj = torch.as_tensor(systems.random_density(n, seed=92)[0], device="cuda")
    ck = torch.as_tensor(systems.random_density(n, seed=93)[0], device="cuda")
    g = plan.jk_param_vjp(d, cj, ck).cpu().numpy()

    def e(env_):
        out = DirectJKPlan(atm, bas, env_, cart=cart).jk(d)
        return float((cj * out[0, 0]).sum() + (ck * out[1, 0]).sum())

    d_shell = int(np.nonzero(bas[:, 1] == 2)[0][0])
    for idx, kind in _slots(bas, [0, 3, d_shell, len(bas) - 1]):
        fd = _fd_env(e, env, idx, kind)
        assert abs(fd - g[idx]) < 1e-6 * max(1.0, abs(fd)), (idx, kind, fd, g[idx])


@pytest.mark.parametrize("basis,cart", [("6-31g*", True), ("cc-pvdz", False)])
def test_int1e_param_vjp_matches_fd(basis, cart):
    import torch
    from gradscf_amd.jk import DirectJKPlan
    spec = M.parse_molecule_spec(WATER, unit="Bohr")
    _, _, atm, bas, env = systems.tables(spec, basis, cart=cart)
    plan = DirectJKPlan(atm, bas, env, cart=cart)
    n = plan.nao
    d = torch.as_tensor(systems.random_density(n, seed=94)[0], device="cuda")
    w = torch.as_tensor(systems.random_density(n, seed=95)[0], device="cuda")
    g = plan.int1e_param_vjp(d, w).cpu().numpy()

    def e(env_):
        s, t, v = DirectJKPlan(atm, bas, env_, cart=cart).int1e()
        return float((d * (t + v)).sum() - (w * s).sum())

    d_shell = int(np.nonzero(bas[:, 1] == 2)[0][0])
    for idx, kind in _slots(bas, [0, 3, d_shell, len(bas) - 1]):
        fd = _fd_env(e, env, idx, kind)
        assert abs(fd - g[idx]) < 1e-6 * max(1.0, abs(fd)), (idx, kind, fd, g[idx])


def test_rhf_basis_gradient_matches_finite_difference():
    """dE/d(exponent) and dE/d(raw coefficient) of the converged RHF energy
    (water 6-31G*, Cartesian) against central differences of SCF energies
    (relative step 1e-5 for exponents, 1e-5 for coefficients)."""
    from gradscf_amd.basis import BasisParameters, prepare_basis
    from gradscf_amd.grad import rhf_basis_gradient
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.packing import pack
    spec = M.parse_molecule_spec(WATER, unit="Bohr")
    topo, par = prepare_basis(spec, "6-31g*", cart=True)
    atm, bas, env = pack(topo, par)
    res = _scf_energy(atm, bas, env, spec, True)
    plan = DirectJKPlan(atm, bas, env, cart=True, cutoff=0.0)
    d_exp, d_coef = rhf_basis_gradient(plan, res.density, res.mo_coeff, res.mo_energy, spec.nelectron // 2,
                                       bas, topo, par)

    def energy(p):
        a, b, e = pack(topo, p)
        return _scf_energy(a, b, e, spec, True).energy

    for s in (0, 3, 5, 6):  # O core s, O p, O d, H s
        for p in range(len(par.exponents[s])):
            h = 1e-5 * par.exponents[s][p]
            vals = []
            for sgn in (1, -1):
                ex = list(par.exponents)
                a = np.array(ex[s], dtype=float)
                a[p] += sgn * h
                ex[s] = a
                vals.append(energy(BasisParameters(tuple(ex), par.coefficients, par.centers, par.nuclear_coords)))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - d_exp[s][p]) < 1e-6 * max(1.0, abs(fd)), ("exp", s, p, fd, d_exp[s][p])
            vals = []
            for sgn in (1, -1):
                cs = list(par.coefficients)
                c = np.array(cs[s], dtype=float)
                c.reshape(len(par.exponents[s]), -1)[p, 0] += sgn * 1e-5
                cs[s] = c
                vals.append(energy(BasisParameters(par.exponents, tuple(cs), par.centers, par.nuclear_coords)))
            fd = (vals[0] - vals[1]) / 2e-5
            ref = np.asarray(d_coef[s]).reshape(len(par.exponents[s]), -1)[p, 0]
            assert abs(fd - ref) < 1e-6 * max(1.0, abs(fd)), ("coef", s, p, fd, ref)
