"""The benchmark workloads at full size on the GPU:

  * C20H42 (all-trans) and (H2O)32 (seed 32, 15 A cube) in 6-31G* Cartesian,
    and the bench workload (H2O)16 (seed 16, 12 A cube) in cc-pVDZ spherical:
    FP64 RHF against the reference-libcint FP64 RHF energies committed in
    tests/golden/scf_energies.json (bar 1e-9 Eh), and the mixed-precision RHF
    (FP32 for quartets with Q_ij Q_kl max|D| < 1e-7) against the same
    energies (bar 1 uEh);
  * (H2O)16 (seed 16, 12 A cube) cc-pVDZ spherical RHF gradients: nuclear
    gradient of all 48 nuclei (translation sum rule, rotational invariance,
    central differences with 1e-4 bohr steps) and dE/d(exponent,
    coefficient) (central differences, steps 1e-5 relative for exponents;
    Richardson extrapolation of steps 1e-4 and 5e-5 for raw coefficients),
    from the FP64 and from the mixed-precision SCF (bar 1e-6 Eh/bohr);
  * the 256-cluster 6-31G* batch in mixed precision: every molecule's energy
    within 1 uEh of the FP64 batch.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from gradscf_amd import molecule as M

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "scf_energies.json"
MIXED = 1e-7


def _golden(key):
    table = json.loads(GOLDEN.read_text())
    assert key in table, f"{key} missing from {GOLDEN.name} (tests/golden/make_golden.py --large)"
    entry = table[key]
    assert entry["converged"]
    return entry


@pytest.mark.parametrize("key,make", [
    ("c20h42_631gs_cart", lambda: M.linear_alkane(20)),
    ("water32_631gs_cart", lambda: M.water_cluster(32, seed=32, box=15.0)),
    ("water16_ccpvdz_sph", lambda: M.water_cluster(16, seed=16, box=12.0)),
])
def test_mixed_precision_workload_energies(key, make):
    from gradscf_amd.scf import RHF
    ref = _golden(key)
    spec = make()
    job = RHF(spec, ref["basis"], cart=ref["cart"], direct_scf_tol=ref["cutoff"])
    assert job.plan.nao == ref["nao"]
    fp64 = job.run()
    assert fp64.converged
    assert abs(fp64.energy - ref["energy"]) < 1e-9, (fp64.energy, ref["energy"])
    mixed = RHF(spec, ref["basis"], cart=ref["cart"], direct_scf_tol=ref["cutoff"], mixed_threshold=MIXED).run()
    assert mixed.converged
    assert abs(mixed.energy - ref["energy"]) < 1e-6, (mixed.energy, ref["energy"])


@pytest.fixture(scope="module")
def water16():
    """Converged FP64 and mixed RHF of (H2O)16 cc-pVDZ with their gradients."""
    from gradscf_amd.basis import prepare_basis
    from gradscf_amd.grad import rhf_basis_gradient, rhf_gradient, shell_nuclei
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.packing import pack
    from gradscf_amd.scf import rhf

    spec = M.water_cluster(16, seed=16, box=12.0)
    topo, par = prepare_basis(spec, "cc-pvdz", cart=False)
    atm, bas, env = pack(topo, par)
    sh = shell_nuclei(par.centers, par.nuclear_coords)
    nocc = spec.nelectron // 2

    def scf(env_, mixed=None):
        p = DirectJKPlan(atm, bas, env_, cart=False, cutoff=1e-12)
        flags = {} if mixed is None else {"mixed_threshold": mixed}
        r = rhf(p, spec.nelectron, 0.0, conv_tol=1e-12, conv_tol_density=1e-10, conv_tol_grad=1e-9,
                jk_flags=flags)
        assert r.converged
        return r

    out = {"spec": spec, "topo": topo, "par": par, "atm": atm, "bas": bas, "env": env, "sh": sh, "scf": scf}
    gplan = DirectJKPlan(atm, bas, env, cart=False, cutoff=0.0)
    for tag, mixed in (("fp64", None), ("mixed", MIXED)):
        r = scf(env, mixed)
        g = rhf_gradient(gplan, r.density, r.mo_coeff, r.mo_energy, nocc, spec.charges, spec.coords_bohr, sh)
        de, dc = rhf_basis_gradient(gplan, r.density, r.mo_coeff, r.mo_energy, nocc, bas, topo, par)
        out[tag] = (r, g, de, dc)
    return out


def _enuc(z, r):
    return sum(z[i] * z[j] / np.linalg.norm(r[i] - r[j]) for i in range(len(z)) for j in range(i))


def test_water16_gradient_invariances(water16):
    spec = water16["spec"]
    r = spec.coords_bohr
    for tag in ("fp64", "mixed"):
        g = water16[tag][1]
        assert g.shape == (48, 3)
        assert np.max(np.abs(g.sum(axis=0))) < 1e-8, tag           # translation
        torque = np.cross(r, g).sum(axis=0)
        assert np.max(np.abs(torque)) < 1e-8, (tag, torque)        # rotation
    g64, gmx = water16["fp64"][1], water16["mixed"][1]
    assert np.max(np.abs(gmx - g64)) < 1e-6
    for a, b in zip(water16["fp64"][2], water16["mixed"][2]):
        assert np.max(np.abs(np.asarray(a) - np.asarray(b))) < 1e-6
    for a, b in zip(water16["fp64"][3], water16["mixed"][3]):
        assert np.max(np.abs(np.asarray(a) - np.asarray(b))) < 1e-6


@pytest.mark.parametrize("atom,axis", [(0, 0), (17, 2), (40, 1)])
def test_water16_nuclear_gradient_finite_difference(water16, atom, axis):
    spec, atm, env, sh = water16["spec"], water16["atm"], water16["env"], water16["sh"]
    rows = [atom] + [len(spec.charges) + i for i in np.nonzero(sh == atom)[0]]
    h = 1e-4
    e = []
    for sgn in (1, -1):
        env_ = np.array(env, copy=True)
        for row in rows:
            env_[int(atm[row, 1]) + axis] += sgn * h
        e.append(water16["scf"](env_).energy)
    cp, cm = spec.coords_bohr.copy(), spec.coords_bohr.copy()
    cp[atom, axis] += h
    cm[atom, axis] -= h
    z = spec.charges.astype(float)
    fd = (e[0] - e[1]) / (2 * h) + (_enuc(z, cp) - _enuc(z, cm)) / (2 * h)
    for tag in ("fp64", "mixed"):
        g = water16[tag][1]
        assert abs(fd - g[atom, axis]) < 1e-6, (tag, atom, axis, fd, g[atom, axis])


@pytest.mark.parametrize("shell,prim,kind", [(0, 3, "exp"), (4, 0, "exp"), (5, 0, "exp"),
                                             (0, 2, "coef"), (2, 1, "coef"), (8, 0, "coef")])
def test_water16_basis_gradient_finite_difference(water16, shell, prim, kind):
    """O 1s/2s (8 primitives, 2 columns), O p, O d, H s: one exponent (central
    difference, step 1e-5 relative) or one raw coefficient of the first
    contraction column (Richardson-extrapolated central differences)."""
    from gradscf_amd.basis import BasisParameters
    from gradscf_amd.packing import pack
    topo, par = water16["topo"], water16["par"]
    vals = []
    if kind == "exp":
        h = 1e-5 * par.exponents[shell][prim]
        for sgn in (1, -1):
            ex = list(par.exponents)
            a = np.array(ex[shell], dtype=float)
            a[prim] += sgn * h
            ex[shell] = a
            _, _, env_ = pack(topo, BasisParameters(tuple(ex), par.coefficients, par.centers, par.nuclear_coords))
            vals.append(water16["scf"](env_).energy)
        fd = (vals[0] - vals[1]) / (2 * h)
    else:
        # The energy is strongly curved in a contraction coefficient (through the
        # column's normalisation): a central difference at 1e-4 is still 7e-6 off
        # for the O 1s column, so two steps (1e-4, 5e-5) are combined by Richardson
        # extrapolation (error O(h^4)).  1e-5 relative steps of small coefficients
        # (O 1s: 0.0278) would sink the difference into the SCF energy's 1e-12 noise.
        c0 = np.asarray(par.coefficients[shell], dtype=float).reshape(len(par.exponents[shell]), -1)
        diffs = []
        for h in (1e-4, 5e-5):
            e = []
            for sgn in (1, -1):
                cs = list(par.coefficients)
                c = c0.copy()
                c[prim, 0] += sgn * h
                cs[shell] = c.reshape(np.asarray(par.coefficients[shell]).shape)
                _, _, env_ = pack(topo, BasisParameters(par.exponents, tuple(cs), par.centers, par.nuclear_coords))
                e.append(water16["scf"](env_).energy)
            diffs.append((e[0] - e[1]) / (2 * h))
        fd = (4.0 * diffs[1] - diffs[0]) / 3.0
    for tag in ("fp64", "mixed"):
        _, _, de, dc = water16[tag]
        if kind == "exp":
            an = float(de[shell][prim])
        else:
            an = float(np.asarray(dc[shell]).reshape(len(par.exponents[shell]), -1)[prim, 0])
        assert abs(fd - an) < 1e-6 * max(1.0, abs(fd)), (tag, shell, prim, kind, fd, an)


def test_mixed_precision_256_cluster_batch():
    """All 256 clusters (n ~ U{1..8} waters, 6-31G*) through the batched SCF in
    mixed precision: converged, every energy within 1 uEh of the FP64 batch."""
    from gradscf_amd.basis import prepare_basis
    from gradscf_amd.batch_scf import rhf_batch
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.packing import pack
    specs = M.batch_water_clusters(256, seed=256, base_seed=1000, nmax=8, box=8.0)
    tabs = [pack(*prepare_basis(s, "6-31g*", cart=True)) for s in specs]
    plan = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
    nel = [s.nelectron for s in specs]
    enuc = [s.nuclear_repulsion() for s in specs]
    ref = rhf_batch(plan, nel, enuc)
    mix = rhf_batch(plan, nel, enuc, mixed_threshold=MIXED)
    assert ref.converged.all() and mix.converged.all()
    assert np.max(np.abs(mix.energies - ref.energies)) < 1e-6
