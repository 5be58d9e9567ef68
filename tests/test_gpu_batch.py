"""Batched plans: many molecules of uneven size in one plan and one launch per
quartet class.  A batched plan must reproduce, molecule by molecule, what a
single-molecule plan of the same tables gives (same pairs, same per-molecule
Schwarz screening, same mixed-precision decisions), and through that the
oracle.  Integer plan statistics (retained pairs and quartets) must add up
exactly; J/K elements agree to 1e-12 (only the order of FP64 atomic additions
differs); the smallest molecules are also checked against the oracle at the
1e-10 FP64 parity bar.
"""
import numpy as np
import pytest

from gradscf_amd import molecule as M
from tests import systems

pytestmark = pytest.mark.gpu


def _mols():
    # uneven sizes, shared generator rules with the 1..8-water batch workload
    return [M.water_monomer(), M.water_cluster(3, seed=1003, box=8.0), M.water_cluster(1, seed=1001, box=8.0),
            M.water_cluster(2, seed=1002, box=8.0)]


def _tables(basis, cart):
    return [systems.tables(spec, basis, cart=cart)[2:] for spec in _mols()]


def _single_plans(tables, cart, cutoff):
    from gradscf_amd.jk import DirectJKPlan
    return [DirectJKPlan(a, b, e, cart=cart, cutoff=cutoff) for a, b, e in tables]


@pytest.mark.parametrize("basis,cart,cutoff", [("6-31g*", True, 0.0), ("6-31g*", True, 1e-12),
                                               ("cc-pvdz", False, 1e-12)])
def test_batch_matches_single_molecule_plans(basis, cart, cutoff):
    import torch
    from gradscf_amd.jk import DirectJKPlan
    tabs = _tables(basis, cart)
    singles = _single_plans(tabs, cart, cutoff)
    plan = DirectJKPlan.batch(tabs, cart=cart, cutoff=cutoff)
    assert plan.nmol == len(tabs)
    assert [int(n) for n in plan.mol_n] == [p.nao for p in singles]
    assert plan.info["npairs"] == sum(p.info["npairs"] for p in singles)
    assert plan.info["nquartets"] == sum(p.info["nquartets"] for p in singles)
    dens = [torch.as_tensor(systems.random_density(p.nao, nb=2, seed=10 + i, symmetric=False), device="cuda")
            for i, p in enumerate(singles)]
    out = plan.jk(plan.pack(dens))
    got = plan.unpack(out)
    for p, d, g in zip(singles, dens, got):
        ref = p.jk(d)
        assert float((g - ref).abs().max()) < 1e-12


def test_batch_matches_oracle(oracle_mod):
    import torch
    from gradscf_amd.jk import DirectJKPlan
    tabs = _tables("6-31g*", True)[:3]
    plan = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
    dens = [systems.random_density(int(n), nb=1, seed=20 + i, symmetric=True) for i, n in enumerate(plan.mol_n)]
    out = plan.jk(plan.pack([torch.as_tensor(d, device="cuda") for d in dens]))
    for (atm, bas, env), d, g in zip(tabs, dens, plan.unpack(out)):
        ref = oracle_mod.direct_jk(atm, bas, env, d, cart=True, cutoff=1e-12, workers=4)
        assert np.max(np.abs(g.cpu().numpy() - ref)) < 1e-10


def test_batch_mixed_precision_matches_single():
    import torch
    from gradscf_amd.jk import DirectJKPlan
    tabs = _tables("6-31g*", True)
    singles = _single_plans(tabs, True, 1e-12)
    plan = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
    dens = [torch.as_tensor(systems.random_density(p.nao, nb=1, seed=30 + i), device="cuda") * 1e-3
            for i, p in enumerate(singles)]
    got = plan.unpack(plan.jk(plan.pack(dens), symmetric=True, mixed_threshold=1e-7))
    for p, d, g in zip(singles, dens, got):
        ref = p.jk(d, symmetric=True, mixed_threshold=1e-7)
        # precision is chosen per wavefront, and the two plans group quartets into
        # waves differently: they may differ by FP32 rounding (2^-24 relative) of
        # quartets whose bound is below 1e-7, a few hundred of them per element
        assert float((g - ref).abs().max()) < 1e-11


@pytest.mark.parametrize("cart", [True, False])
def test_batch_int1e_and_schwarz_match_single(cart):
    from gradscf_amd.jk import DirectJKPlan
    tabs = _tables("cc-pvdz", cart)
    singles = _single_plans(tabs, cart, 1e-12)
    plan = DirectJKPlan.batch(tabs, cart=cart, cutoff=1e-12)
    s, t, v = plan.int1e()
    for p, a, b, c in zip(singles, plan.unpack(s), plan.unpack(t), plan.unpack(v)):
        rs, rt, rv = p.int1e()
        assert float((a - rs).abs().max()) < 1e-13
        assert float((b - rt).abs().max()) < 1e-13
        assert float((c - rv).abs().max()) < 1e-12
    for p, q in zip(singles, plan.schwarz()):
        assert np.array_equal(p.schwarz(), q)


def test_batch_geometry_vjp_matches_single():
    import torch
    from gradscf_amd.jk import DirectJKPlan
    tabs = _tables("6-31g*", True)[:2]
    singles = _single_plans(tabs, True, 0.0)
    plan = DirectJKPlan.batch(tabs, cart=True, cutoff=0.0)
    ds, cjs, cks = [], [], []
    for i, p in enumerate(singles):
        ds.append(torch.as_tensor(systems.random_density(p.nao, seed=40 + i)[0], device="cuda"))
        cjs.append(torch.as_tensor(systems.random_density(p.nao, seed=50 + i)[0], device="cuda"))
        cks.append(torch.as_tensor(systems.random_density(p.nao, seed=60 + i)[0], device="cuda"))
    g = plan.jk_geometry_vjp(plan.pack(ds), plan.pack(cjs), plan.pack(cks))
    o = 0
    for p, d, cj, ck in zip(singles, ds, cjs, cks):
        ref = p.jk_geometry_vjp(d, cj, ck)
        assert float((g[o:o + len(p.atm)] - ref).abs().max()) < 1e-11
        o += len(p.atm)


def _batch_rhf_setup(basis="6-31g*"):
    from gradscf_amd.jk import DirectJKPlan
    # the last molecule is the geometry of the reference's published RHF energies
    specs = _mols() + [M.parse_molecule_spec("O 0 0 0; H 0 1.43 1.11; H 0 -1.43 1.11", unit="Bohr")]
    tabs = [systems.tables(s, basis, cart=True)[2:] for s in specs]
    plan = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
    return specs, tabs, plan


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_rhf_matches_single_molecule_rhf(use_graph):
    """Batched SCF energies equal independent single-molecule runs (1e-9 Eh)."""
    from gradscf_amd.batch_scf import rhf_batch
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.scf import rhf
    specs, tabs, plan = _batch_rhf_setup()
    res = rhf_batch(plan, [s.nelectron for s in specs], [s.nuclear_repulsion() for s in specs],
                    use_graph=use_graph)
    assert res.converged.all()
    for spec, (a, b, e), energy in zip(specs, tabs, res.energies):
        ref = rhf(DirectJKPlan(a, b, e, cart=True, cutoff=1e-12), spec.nelectron, spec.nuclear_repulsion())
        assert ref.converged
        assert abs(energy - ref.energy) < 1e-9
    # reference-published RHF/6-31G* energy (examples/geometry/water_correlated_forces.py:84-87)
    assert abs(res.energies[-1] - (-76.010492818245)) < 1e-9


def test_batch_rhf_mixed_within_micro_hartree():
    from gradscf_amd.batch_scf import rhf_batch
    specs, tabs, plan = _batch_rhf_setup()
    args = ([s.nelectron for s in specs], [s.nuclear_repulsion() for s in specs])
    ref = rhf_batch(plan, *args, use_graph=False)
    mixed = rhf_batch(plan, *args, mixed_threshold=1e-7, use_graph=True)
    assert mixed.converged.all()
    assert np.max(np.abs(mixed.energies - ref.energies)) < 1e-6


def test_graph_replay_matches_eager_every_time():
    """The captured J/K graph, replayed several times with eager work and
    synchronisations in between, equals the eager build on every replay."""
    import torch
    from gradscf_amd.batch_scf import _GraphJK
    specs, tabs, plan = _batch_rhf_setup()
    g = _GraphJK(plan, None)
    for i in range(4):
        dens = plan.pack([torch.as_tensor(systems.random_density(int(n), seed=70 + 10 * i + m)[0], device="cuda")
                          for m, n in enumerate(plan.mol_n)])
        torch.cuda.synchronize()
        j, k = g(dens)
        eager = plan.jk(dens[None], symmetric=True)
        torch.cuda.synchronize()
        assert float((j - eager[0, 0]).abs().max()) < 1e-12
        assert float((k - eager[1, 0]).abs().max()) < 1e-12
        _ = (dens * 2.0).sum().item()  # eager work between replays


def test_grouped_batched_eigh_matches_torch():
    """linalg.BatchedEigh (rocSOLVER's batched solver per AO-count group, the
    groups on concurrent streams) against torch.linalg.eigh of every
    molecule's block: eigenvalues to 1e-12 relative, eigenvectors through the
    residual and orthonormality, identity in the padding, and molecules
    outside the active set left untouched; a second call with other blocks
    reuses the handles and streams."""
    import torch
    from gradscf_amd.linalg import BatchedEigh
    rng = np.random.default_rng(5)
    sizes = [19 * int(m) for m in rng.integers(1, 9, size=40)] + [152, 19]
    k, N = len(sizes), max(sizes)
    eig = BatchedEigh(sizes, torch.device("cuda", 0))
    c_prev = None
    for rep in range(2):
        a = torch.zeros((k, N, N), dtype=torch.float64)
        for m, n in enumerate(sizes):
            x = rng.normal(size=(n, n))
            a[m, :n, :n] = torch.as_tensor(x + x.T)
        a = a.cuda()
        active = np.ones(k, dtype=bool)
        if rep:
            active[::3] = False
        e, c = eig(a, c_out=None if c_prev is None else c_prev.clone(), active_h=active)
        torch.cuda.synchronize()
        for m, n in enumerate(sizes):
            if not active[m]:
                assert torch.equal(c[m], c_prev[m])
                continue
            blk = a[m, :n, :n]
            e0 = torch.linalg.eigh(blk)[0]
            assert float(((e[m, :n] - e0).abs() / e0.abs().max()).max()) < 1e-12
            v = c[m, :n, :n]
            assert float((blk @ v - v * e[m, :n][None, :]).abs().max()) < 1e-10 * float(e0.abs().max())
            assert float((v.T @ v - torch.eye(n, dtype=torch.float64, device="cuda")).abs().max()) < 1e-12
            if n < N:
                assert float((c[m, n:, n:] - torch.eye(N - n, dtype=torch.float64, device="cuda")).abs().max()) == 0
                assert float(c[m, :n, n:].abs().max()) == 0 and float(c[m, n:, :n].abs().max()) == 0
        c_prev = c
    eig.check()
