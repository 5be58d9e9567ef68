"""Pins of the oracle (CPU).  Each check is chosen so that a plausible slip in the
restatement (a dropped term, a wrong sign or index, a transposed operand)
fails it:
  * Boys F_m against 40-digit mpmath, closed forms and the recursion identity;
  * Rys roots/weights against the moment identities sum w s^k = F_k (k < 2n)
    and against the reference libcint CINTrys_roots (when oracle/_ref is built);
  * (ss|ss) closed form, 8-fold permutational symmetry of ERI blocks;
  * ERI / int2e_ip1 / one-electron blocks against the reference libcint build;
  * J/K against the brute-force full-tensor contraction (the reference's own
    J/K test oracle, tests/integrals/test_native_direct_parallel.py:45-60);
  * RHF energies against the reference's published values
    (examples/geometry/water_correlated_forces.py:25-30,84-87);
  * the committed golden fixtures (tests/golden/).
"""
import math

import numpy as np
import pytest

from gradscf_amd import molecule as M
from tests import systems

ref_only = pytest.mark.skipif("not __import__('oracle').have_ref()", reason="oracle/_ref not built")


def test_boys_vs_mpmath_and_closed_forms(oracle_mod):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    for x in [0.0, 1e-12, 1e-3, 0.5, 3.0, 12.0, 29.9, 44.99, 45.01, 60.0, 150.0]:
        f = oracle_mod.boys(16, x)
        for m in range(17):
            exact = mpmath.hyp1f1(m + 0.5, m + 1.5, -x) / (2 * m + 1)
            assert abs(f[m] - float(exact)) <= 4e-16 * float(exact)
        if x > 0:
            assert abs(f[0] - 0.5 * math.sqrt(math.pi / x) * math.erf(math.sqrt(x))) < 1e-15
            for m in range(16):  # downward recursion identity
                assert abs((2 * m + 1) * f[m] - (2 * x * f[m + 1] + math.exp(-x))) < 1e-14 * (2 * m + 1) * f[m]
    f0 = oracle_mod.boys(10, 0.0)
    assert np.allclose(f0, [1 / (2 * m + 1) for m in range(11)], rtol=0, atol=1e-16)


@pytest.mark.parametrize("n", range(1, 8))
def test_rys_moment_identities(oracle_mod, n):
    for x in [0.0, 1e-9, 0.7, 4.0, 11.0, 27.0, 55.0, 89.0, 91.0, 300.0, 5e4]:
        s, w = oracle_mod.rys_roots(n, x)
        assert np.all(np.diff(s) > 0) and np.all((s > 0) & (s < 1)) and np.all(w > 0)
        f = oracle_mod.boys(2 * n - 1, x)
        for k in range(2 * n):
            assert abs(np.sum(w * s ** k) - f[k]) <= 5e-15 * f[k]


@ref_only
@pytest.mark.parametrize("n", range(1, 8))
def test_rys_vs_libcint(oracle_mod, n):
    # libcint's own fitted roots are accurate to ~1e-11 relative for 6-7 roots;
    # the oracle satisfies the moment identities to 1e-15 (test above).
    tol = 1e-12 if n <= 5 else 5e-10
    for x in [0.0, 0.3, 2.0, 7.5, 15.0, 33.0, 49.0, 80.0, 200.0]:
        s, w = oracle_mod.rys_roots(n, x)
        u, wl = oracle_mod.libcint_rys_roots(n, x)
        assert np.max(np.abs(s - u / (1 + u)) / s) < tol
        assert np.max(np.abs(w - wl) / wl) < tol * 10


def test_ssss_closed_form(oracle_mod):
    """Single normalised s primitives on two centres: (ss|ss) = 2 pi^{5/2}/(pq sqrt(p+q)) K K F0(T)."""
    spec = M.parse_molecule_spec("H 0 0 0; H 0 0 1.3", unit="Bohr")
    basis = {"H": [[0, [0.8, 1.0]]]}
    _, _, atm, bas, env = systems.tables(spec, basis)
    a = 0.8
    N = (2 * a / math.pi) ** 0.75  # normalised s primitive
    for q in [(0, 0, 0, 0), (0, 1, 0, 1), (0, 0, 1, 1), (0, 1, 1, 0)]:
        R = [np.array([0, 0, 0.0]), np.array([0, 0, 1.3])]
        A, B, C, D = (R[i] for i in q)
        p = qq = 2 * a
        P, Q = (A + B) / 2, (C + D) / 2
        kab = math.exp(-a * a / p * np.sum((A - B) ** 2))
        kcd = math.exp(-a * a / qq * np.sum((C - D) ** 2))
        T = p * qq / (p + qq) * np.sum((P - Q) ** 2)
        f0 = 1.0 if T == 0 else 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
        exact = N ** 4 * 2 * math.pi ** 2.5 / (p * qq * math.sqrt(p + qq)) * kab * kcd * f0
        got = oracle_mod.eri_block(q, atm, bas, env, True)[0, 0, 0, 0]
        assert abs(got - exact) < 1e-14 * exact


def test_eri_permutational_symmetry(oracle_mod):
    _, _, atm, bas, env = systems.tables(systems.water(), "6-31g*", cart=True)
    rng = np.random.default_rng(5)
    for _ in range(12):
        i, j, k, l = (int(v) for v in rng.integers(0, len(bas), 4))
        b = oracle_mod.eri_block((i, j, k, l), atm, bas, env, True)
        assert np.allclose(oracle_mod.eri_block((j, i, k, l), atm, bas, env, True), b.transpose(1, 0, 2, 3),
                           atol=1e-14)
        assert np.allclose(oracle_mod.eri_block((k, l, i, j), atm, bas, env, True), b.transpose(2, 3, 0, 1),
                           atol=1e-14)
        assert np.allclose(oracle_mod.eri_block((i, j, l, k), atm, bas, env, True), b.transpose(0, 1, 3, 2),
                           atol=1e-14)


@ref_only
@pytest.mark.parametrize("basis,cart", [("6-31g*", True), ("cc-pvdz", True), ("cc-pvdz", False)])
def test_blocks_vs_libcint(oracle_mod, basis, cart):
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), basis, cart=cart)
    rng = np.random.default_rng(9)
    for _ in range(25):
        q = tuple(int(v) for v in rng.integers(0, len(bas), 4))
        a = oracle_mod.eri_block(q, atm, bas, env, cart)
        b = oracle_mod.libcint_block(q, atm, bas, env, cart)
        assert np.max(np.abs(a - b)) < 3e-13
        a1 = oracle_mod.eri_block(q, atm, bas, env, cart, deriv=1)
        b1 = oracle_mod.libcint_block(q, atm, bas, env, cart, "int2e_ip1")
        assert np.max(np.abs(a1 - b1)) < 3e-13


@ref_only
@pytest.mark.parametrize("cart", [True, False])
def test_int1e_vs_libcint(oracle_mod, cart):
    _, _, atm, bas, env = systems.tables(systems.water(), "cc-pvdz", cart=cart)
    mine = oracle_mod.int1e(atm, bas, env, cart)
    ref = oracle_mod.libcint_int1e(atm, bas, env, cart)
    for a, b in zip(mine, ref):
        assert np.max(np.abs(a - b)) < 1e-12


@pytest.mark.parametrize("cart", [True, False])
def test_jk_vs_full_tensor(oracle_mod, cart):
    _, _, atm, bas, env = systems.tables(systems.water(), "cc-pvdz", cart=cart)
    n = oracle_mod.nao(bas, cart)
    d = systems.random_density(n, nb=2, seed=3, symmetric=False)
    eri = oracle_mod.full_eri(atm, bas, env, cart)
    j = np.einsum("pqrs,brs->bpq", eri, d)
    k = np.einsum("prqs,brs->bpq", eri, d)
    for workers in (1, 4):
        out = oracle_mod.direct_jk(atm, bas, env, d, cart=cart, workers=workers)
        assert np.max(np.abs(out[0] - j)) < 2e-13
        assert np.max(np.abs(out[1] - k)) < 2e-13


@ref_only
def test_driver_on_libcint_matches_own_rys(oracle_mod):
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "6-31g*", cart=True)
    n = oracle_mod.nao(bas, True)
    d = systems.random_density(n, seed=4)
    own = oracle_mod.direct_jk(atm, bas, env, d, cart=True, cutoff=1e-12, workers=4)
    with oracle_mod.backend("libcint"):
        ref = oracle_mod.direct_jk(atm, bas, env, d, cart=True, cutoff=1e-12, workers=4)
    assert np.max(np.abs(own - ref)) < 1e-12


@pytest.mark.parametrize("basis,energy", [("6-31g*", -76.010492818245), ("sto-3g", -74.963082931333)])
def test_published_rhf_energies(oracle_mod, basis, energy):
    from oracle import scf
    spec = M.parse_molecule_spec("O 0 0 0; H 0 1.43 1.11; H 0 -1.43 1.11", unit="Bohr")
    _, _, atm, bas, env = systems.tables(spec, basis, cart=True)
    s, t, v = oracle_mod.int1e(atm, bas, env, True)

    def jk(d):
        o = oracle_mod.direct_jk(atm, bas, env, d, cart=True, workers=4)
        return o[0, 0], o[1, 0]

    r = scf.rhf(t + v, s, spec.nelectron, spec.nuclear_repulsion(), jk)
    assert r["converged"]
    assert abs(r["energy"] - energy) < 1e-9


def test_oracle_matches_golden_fixtures(oracle_mod):
    """The own-Rys oracle reproduces the fixtures made on the reference libcint."""
    import json
    from pathlib import Path
    g = Path(__file__).resolve().parent / "golden"
    f = np.load(g / "water_631g_jk.npz")
    _, _, atm, bas, env = systems.tables(M.water_monomer(), "6-31g", cart=True)
    out = oracle_mod.direct_jk(atm, bas, env, f["density"], cart=True)
    assert np.max(np.abs(out[0] - f["J"])) < 1e-12 and np.max(np.abs(out[1] - f["K"])) < 1e-12
    f = np.load(g / "dimer_ccpvdz_sph_jk.npz")
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=False)
    out = oracle_mod.direct_jk(atm, bas, env, f["density"], cart=False, cutoff=1e-12, workers=4)
    assert np.max(np.abs(out[0, 0] - f["J"])) < 1e-11 and np.max(np.abs(out[1, 0] - f["K"])) < 1e-11
    e = json.loads((g / "scf_energies.json").read_text())["water_631g_cart"]["energy"]
    from oracle import scf
    spec = M.water_monomer()
    _, _, atm, bas, env = systems.tables(spec, "6-31g", cart=True)
    s, t, v = oracle_mod.int1e(atm, bas, env, True)
    r = scf.rhf(t + v, s, spec.nelectron, spec.nuclear_repulsion(),
                lambda d: tuple(oracle_mod.direct_jk(atm, bas, env, d, cart=True)[:, 0]))
    assert abs(r["energy"] - e) < 1e-10


def test_oracle_geometry_vjp_pins(oracle_mod):
    """Coordinate VJP: translation sum rule, central finite differences of
    E = tr(D J)/2 - tr(D K)/4, and (with _ref) the same loop on libcint int2e_ip1."""
    _, _, atm, bas, env = systems.tables(systems.water(), "6-31g", cart=True)
    n = 13
    rng = np.random.default_rng(0)
    d = rng.normal(size=(n, n))
    d = d + d.T
    g = oracle_mod.jk_geometry_vjp(atm, bas, env, d, 0.5 * d, -0.25 * d, cart=True)
    assert np.abs(g.sum(axis=0)).max() < 1e-11

    def energy(e):
        o = oracle_mod.direct_jk(atm, bas, e, d, cart=True)
        return 0.5 * np.sum(d * o[0, 0]) - 0.25 * np.sum(d * o[1, 0])

    h = 1e-4
    for row in (3, 5, 10):
        for x in range(3):
            e1, e2 = env.copy(), env.copy()
            e1[atm[row, 1] + x] += h
            e2[atm[row, 1] + x] -= h
            fd = (energy(e1) - energy(e2)) / (2 * h)
            assert abs(fd - g[row, x]) < 3e-6 * max(1.0, abs(fd))
    if oracle_mod.have_ref():
        cj, ck = rng.normal(size=(n, n)), rng.normal(size=(n, n))
        own = oracle_mod.jk_geometry_vjp(atm, bas, env, d, cj, ck, cart=True)
        with oracle_mod.backend("libcint"):
            ref = oracle_mod.jk_geometry_vjp(atm, bas, env, d, cj, ck, cart=True)
        assert np.max(np.abs(own - ref)) < 1e-10


# ---- Schwarz bounds and the screened quartet loop -------------------------------------------
# Reference semantics (csrc/coulomb_shells.h:60-77, 98-110): Q_ij = sqrt(max over AO pairs
# (p in i, q in j) of |(pq|pq)|); a canonical quartet is skipped when Q_ij*Q_kl < cutoff (the
# pair prune Q_ij*max(Q) < cutoff only removes quartets the skip removes too).  The pins below
# rebuild Q from ERI blocks in the test, bound every block by it, and compare the screened
# driver with the full s1 tensor masked by exactly the quartets the rule keeps.

def _block_fn(oracle_mod, which):
    if which == "libcint":
        return lambda q, atm, bas, env, cart: oracle_mod.libcint_block(q, atm, bas, env, cart)
    return lambda q, atm, bas, env, cart: oracle_mod.eri_block(q, atm, bas, env, cart)


def _schwarz_from_blocks(block, atm, bas, env, cart):
    nb = len(bas)
    q = np.zeros((nb, nb))
    for i in range(nb):
        for j in range(i + 1):
            b = block((i, j, i, j), atm, bas, env, cart)
            di, dj = b.shape[0], b.shape[1]
            diag = np.abs(b.reshape(di * dj, di * dj).diagonal())
            q[i, j] = q[j, i] = math.sqrt(diag.max())
    return q


def _canonical_quartets(nb):
    for i in range(nb):
        for j in range(i + 1):
            for k in range(i + 1):
                for l in range(k + 1 if k < i else j + 1):
                    yield i, j, k, l


def _full_tensor_and_mask(block, atm, bas, env, cart, qmat, cutoff):
    """Full s1 tensor from canonical blocks (8 permutations each) and the 0/1 mask of the AO
    quartets whose shell quartet survives Q_ij*Q_kl >= cutoff."""
    ao = np.asarray(__import__("oracle").ao_offsets(bas, cart))
    n = int(ao[-1])
    eri = np.zeros((n, n, n, n))
    keep = np.zeros((n, n, n, n))
    dropped = 0
    for i, j, k, l in _canonical_quartets(len(bas)):
        b = block((i, j, k, l), atm, bas, env, cart)
        kept = qmat[i, j] * qmat[k, l] >= cutoff
        dropped += not kept
        s = [slice(ao[x], ao[x + 1]) for x in (i, j, k, l)]
        for perm in ((0, 1, 2, 3), (1, 0, 2, 3), (0, 1, 3, 2), (1, 0, 3, 2),
                     (2, 3, 0, 1), (3, 2, 0, 1), (2, 3, 1, 0), (3, 2, 1, 0)):
            idx = tuple(s[p] for p in perm)
            eri[idx] = b.transpose(perm)
            keep[idx] = 1.0 if kept else 0.0
    return eri, keep, dropped


@pytest.mark.parametrize("which", ["oracle", pytest.param("libcint", marks=ref_only)])
@pytest.mark.parametrize("cart", [True, False])
def test_schwarz_matches_block_diagonal_definition(oracle_mod, which, cart):
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=cart)
    mine = oracle_mod.schwarz(atm, bas, env, cart)
    want = _schwarz_from_blocks(_block_fn(oracle_mod, which), atm, bas, env, cart)
    low = np.tril_indices(len(bas))
    # relative 1e-13; the absolute floor covers pairs with Q ~ 1e-11 ((pq|pq) ~ 1e-22), where
    # libcint's primitive-pair cut changes the block in its tenth digit
    assert np.all(np.abs(mine[low] - want[low]) <= 1e-13 * want[low] + 1e-18)


@pytest.mark.parametrize("cart", [True, False])
def test_schwarz_bounds_every_block(oracle_mod, cart):
    """Cauchy-Schwarz: |(ij|kl)| <= Q_ij Q_kl for every element, with equality on (ij|ij)."""
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=cart)
    q = oracle_mod.schwarz(atm, bas, env, cart)
    q = np.tril(q) + np.tril(q, -1).T
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(300):
        i, j, k, l = (int(v) for v in rng.integers(0, len(bas), 4))
        b = oracle_mod.eri_block((i, j, k, l), atm, bas, env, cart)
        worst = max(worst, float(np.abs(b).max() / (q[i, j] * q[k, l])))
    assert worst <= 1.0 + 1e-12
    for i in range(len(bas)):
        b = oracle_mod.eri_block((i, 0, i, 0), atm, bas, env, cart)
        assert abs(np.abs(b).max() - q[i, 0] ** 2) <= 1e-13 * q[i, 0] ** 2


@pytest.mark.parametrize("cart", [True, False])
@pytest.mark.parametrize("cutoff", [1e-12, 1e-7])
def test_screened_jk_equals_masked_full_tensor(oracle_mod, cart, cutoff):
    """The screened driver keeps exactly the quartets with Q_ij Q_kl >= cutoff (masked einsum to
    1e-13), and at 1e-12 differs from the unscreened full-tensor J/K by <= 1e-10."""
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=cart)
    block = _block_fn(oracle_mod, "libcint" if oracle_mod.have_ref() else "oracle")
    qmat = _schwarz_from_blocks(block, atm, bas, env, cart)
    eri, keep, dropped = _full_tensor_and_mask(block, atm, bas, env, cart, qmat, cutoff)
    assert dropped > 1000  # the screening branch really removes quartets here
    n = eri.shape[0]
    d = systems.random_density(n, nb=2, seed=21, symmetric=False)
    masked = eri * keep
    jm = np.einsum("pqrs,brs->bpq", masked, d)
    km = np.einsum("prqs,brs->bpq", masked, d)
    for backend in (["oracle", "libcint"] if oracle_mod.have_ref() else ["oracle"]):
        with oracle_mod.backend(backend):
            out = oracle_mod.direct_jk(atm, bas, env, d, cart=cart, cutoff=cutoff, workers=4)
        assert np.max(np.abs(out[0] - jm)) < 1e-12
        assert np.max(np.abs(out[1] - km)) < 1e-12
        if cutoff == 1e-12:
            jf = np.einsum("pqrs,brs->bpq", eri, d)
            kf = np.einsum("prqs,brs->bpq", eri, d)
            assert np.max(np.abs(out[0] - jf)) < 1e-10
            assert np.max(np.abs(out[1] - kf)) < 1e-10


def _df_tables(aux="cc-pvdz-jkfit", cart=True, spec=None):
    """Orbital (6-31G*) and auxiliary shells of water in one table set, like the
    reference's make_auxiliary_plan (orbital shells first)."""
    from gradscf_amd import molecule as M
    from gradscf_amd.basis import BasisParameters, BasisTopology, prepare_basis
    from gradscf_amd.packing import pack
    spec = spec or M.water_monomer()
    top, p = prepare_basis(spec, "6-31g*", cart=cart)
    at, a = prepare_basis(spec, aux, cart=cart)
    comb = BasisTopology(top.angular_momenta + at.angular_momenta, top.primitive_counts + at.primitive_counts,
                         top.contraction_counts + at.contraction_counts, top.nuclear_charges, cart)
    par = BasisParameters(p.exponents + a.exponents, p.coefficients + a.coefficients,
                          np.concatenate([p.centers, a.centers]), p.nuclear_coords)
    atm, bas, env = pack(comb, par)
    return atm, bas, env, top.nshell


def _with_unit(atm, bas, env, atom_row):
    """Append the unit function: one s primitive of exponent 0 with coefficient
    1/fac_sp(0) = sqrt(4 pi) (cancelling the s factor the four-centre driver
    applies), on the atm row of the shell it is paired with."""
    env = np.concatenate([env, [0.0, np.sqrt(4.0 * np.pi)]])
    row = np.array([[atom_row, 0, 1, 1, 0, len(env) - 2, len(env) - 1, 0]], dtype=np.int32)
    return atm, np.concatenate([bas, row]), env


@ref_only
def test_three_and_two_centre_integrals_pin_the_unit_function(oracle_mod):
    """libcint's int3c2e (ab|P) equals the oracle's own four-centre Rys ERI
    (ab|P u) and int2c2e (P|Q) equals (P u|Q u), u the unit function: the
    three- and two-centre integrals the GPU evaluates through its four-centre
    kernels, pinned on s..f auxiliary shells (Cartesian)."""
    atm, bas, env, split = _df_tables(cart=True)
    rng = np.random.default_rng(41)
    ls = bas[:, 1]
    aux = np.arange(split, len(bas))
    for k in [int(aux[np.nonzero(ls[aux] == l)[0][0]]) for l in range(4)]:
        a2, b2, e2 = _with_unit(atm, bas, env, int(bas[k, 0]))
        u = len(b2) - 1
        for _ in range(6):
            i, j = (int(x) for x in rng.integers(0, split, 2))
            ref = oracle_mod.libcint_aux_block((i, j, k), atm, bas, env, True)
            own = oracle_mod.eri_block((i, j, k, u), a2, b2, e2, True)[..., 0]
            assert np.max(np.abs(own - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
        for kk in [int(aux[np.nonzero(ls[aux] == l)[0][-1]]) for l in range(4)]:
            a3, b3, e3 = _with_unit(a2, b2, e2, int(bas[kk, 0]))
            ref = oracle_mod.libcint_aux_block((k, kk), atm, bas, env, True)
            own = oracle_mod.eri_block((k, u, kk, len(b3) - 1), a3, b3, e3, True)[:, 0, :, 0]
            assert np.max(np.abs(own - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


@ref_only
def test_density_fitting_oracle_layouts(oracle_mod):
    """The restated compact layouts 2 and 3: the metric is symmetric positive
    definite, and (ab|P) contracted with the inverse metric reproduces the
    fitted Coulomb energy below the exact one (the Coulomb fitting error
    (rho - rho~|rho - rho~) is non-negative) and within 1e-3 of it."""
    atm, bas, env, split = _df_tables(cart=False)
    metric = oracle_mod.density_fitting(atm, bas, env, layout=2, cart=False, split=split)
    three = oracle_mod.density_fitting(atm, bas, env, layout=3, cart=False, split=split)
    assert np.array_equal(metric, metric.T)
    assert np.linalg.eigvalsh(metric).min() > 0
    n = int(round((np.sqrt(8 * three.shape[1] + 1) - 1) / 2))
    d = systems.random_density(n, seed=42)[0]
    d = d @ d.T / n
    i, j = np.tril_indices(n)
    w = np.where(i == j, 1.0, 2.0) * d[i, j]
    rho = three @ w                              # (P|rho)
    e_fit = rho @ np.linalg.solve(metric, rho)   # (rho~|rho~) = (rho|P) V^-1 (P|rho)
    jk = oracle_mod.direct_jk(atm, bas[:split], env, d, cart=False)
    e_exact = float(np.sum(d * jk[0, 0]))        # (rho|rho)
    assert e_fit <= e_exact + 1e-12
    assert e_exact - e_fit < 1e-3 * e_exact
