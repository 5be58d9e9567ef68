"""In-core integrals (the plan-resident form of the reference's packed-ERI
route, csrc/packed_eri.cc + csrc/packed_jk.cc): after DirectJKPlan.incore()
every build digests stored contracted blocks.  The stored blocks are the
evaluated ones, so an in-core build equals the evaluating build to rounding
of the FP64 atomic order (1e-12) and the oracle to the 1e-10 parity bar, for
symmetric and nonsymmetric densities, J or K alone, single and batched plans,
graph replays and whole SCF runs.
"""
import numpy as np
import pytest

from gradscf_amd import molecule as M
from tests import systems

pytestmark = pytest.mark.gpu


def _cuda(x):
    import torch
    return torch.as_tensor(x, device="cuda")


@pytest.mark.parametrize("basis,cart,cutoff", [("6-31g*", True, 0.0), ("cc-pvdz", False, 1e-12)])
def test_incore_matches_evaluating_build_and_oracle(oracle_mod, basis, cart, cutoff):
    from gradscf_amd.jk import DirectJKPlan
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), basis, cart=cart)
    plan = DirectJKPlan(atm, bas, env, cart=cart, cutoff=cutoff)
    n = plan.nao
    d = systems.random_density(n, nb=2, seed=31, symmetric=False)
    ref_direct = plan.jk(_cuda(d)).cpu().numpy()
    nbytes = plan.incore_bytes()
    assert nbytes > 0 and nbytes % 8 == 0
    held = plan.device_bytes()
    plan.incore(True)
    plan.incore(True)  # idempotent
    assert plan.device_bytes() == held + nbytes
    got = plan.jk(_cuda(d)).cpu().numpy()
    assert np.max(np.abs(got - ref_direct)) < 1e-12
    ref = oracle_mod.direct_jk(atm, bas, env, d, cart=cart, cutoff=cutoff, workers=2)
    assert np.max(np.abs(got - ref)) < 1e-10
    # one operator at a time, symmetric density
    ds = _cuda(systems.random_density(n, nb=1, seed=32, symmetric=True))
    j_only = plan.jk(ds, with_k=False)
    k_only = plan.jk(ds, with_j=False)
    both = plan.jk(ds)
    assert float((j_only[0] - both[0]).abs().max()) < 1e-12
    assert float((k_only[0] - both[1]).abs().max()) < 1e-12
    # the mixed-precision flag leaves stored (FP64) blocks exact
    mixed = plan.jk(ds, mixed_threshold=1e-7)
    assert float((mixed - both).abs().max()) < 1e-12
    plan.incore(False)
    again = plan.jk(_cuda(d)).cpu().numpy()
    assert np.max(np.abs(again - ref_direct)) < 1e-12


def test_incore_batch_matches_evaluating_build():
    import torch
    from gradscf_amd.jk import DirectJKPlan
    specs = [M.water_monomer(), M.water_cluster(3, seed=1003, box=8.0), M.water_cluster(2, seed=1002, box=8.0)]
    tabs = [systems.tables(s, "6-31g*", cart=True)[2:] for s in specs]
    plan = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
    dens = plan.pack([torch.as_tensor(systems.random_density(int(n), nb=2, seed=40 + m, symmetric=False),
                                      device="cuda") for m, n in enumerate(plan.mol_n)])
    ref = plan.jk(dens)
    plan.incore(True)
    got = plan.jk(dens)
    assert float((got - ref).abs().max()) < 1e-12


def test_incore_graph_replay_matches_eager_every_time():
    """A J/K graph captured on an in-core plan, replayed with eager work and
    synchronisations in between, equals the eager build on every replay."""
    import torch
    from gradscf_amd.batch_scf import _GraphJK
    from gradscf_amd.jk import DirectJKPlan
    specs = [M.water_monomer(), M.water_cluster(2, seed=1002, box=8.0)]
    tabs = [systems.tables(s, "6-31g*", cart=True)[2:] for s in specs]
    plan = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
    plan.incore(True)
    g = _GraphJK(plan, None)
    for i in range(3):
        dens = plan.pack([torch.as_tensor(systems.random_density(int(n), seed=80 + 10 * i + m)[0], device="cuda")
                          for m, n in enumerate(plan.mol_n)])
        torch.cuda.synchronize()
        j, k = g(dens)
        eager = plan.jk(dens[None], symmetric=True)
        torch.cuda.synchronize()
        assert float((j - eager[0, 0]).abs().max()) < 1e-12
        assert float((k - eager[1, 0]).abs().max()) < 1e-12
        _ = (dens * 2.0).sum().item()


def test_incore_scf_matches_direct_scf():
    from gradscf_amd.batch_scf import rhf_batch
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.scf import rhf
    spec = M.parse_molecule_spec("O 0 0 0; H 0 1.43 1.11; H 0 -1.43 1.11", unit="Bohr")
    _, _, atm, bas, env = systems.tables(spec, "6-31g*", cart=True)
    plan = DirectJKPlan(atm, bas, env, cart=True, cutoff=1e-12)
    direct = rhf(plan, spec.nelectron, spec.nuclear_repulsion())
    incore = rhf(plan, spec.nelectron, spec.nuclear_repulsion(), incore=True)
    assert direct.converged and incore.converged
    assert abs(incore.energy - direct.energy) < 1e-10
    # reference-published RHF/6-31G* energy (examples/geometry/water_correlated_forces.py:84-87)
    assert abs(incore.energy - (-76.010492818245)) < 1e-9
    assert not plan.incore_enabled
    specs = [M.water_monomer(), M.water_cluster(2, seed=1002, box=8.0), spec]
    tabs = [systems.tables(s, "6-31g*", cart=True)[2:] for s in specs]
    bplan = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
    args = ([s.nelectron for s in specs], [s.nuclear_repulsion() for s in specs])
    ref = rhf_batch(bplan, *args)
    got = rhf_batch(bplan, *args, incore=True)
    assert got.converged.all()
    assert np.max(np.abs(got.energies - ref.energies)) < 1e-10


def test_incore_with_general_contraction_matches_evaluating_build():
    """A plan with native general contraction keeps its single-column classes
    in core and evaluates the multi-column patterns in the same build."""
    from gradscf_amd import autotune as at
    from gradscf_amd.jk import DirectJKPlan
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=False)
    before = at.get_options()
    try:
        at.set_options(general_contraction=1)
        plan = DirectJKPlan(atm, bas, env, cart=False, cutoff=1e-12)
    finally:
        at.set_options(**before)
    d = _cuda(systems.random_density(plan.nao, nb=1, seed=33, symmetric=False))
    ref = plan.jk(d)
    assert plan.incore_bytes() > 0
    plan.incore(True)
    got = plan.jk(d)
    assert float((got - ref).abs().max()) < 1e-12


def test_full_256_cluster_batch_properties():
    """The 256-cluster 6-31G* batch at full size (1.07e8 quartets): in-core
    builds equal evaluating builds, J/K are linear in the density, and every
    molecule's J and K of a symmetric density are symmetric."""
    import torch
    from gradscf_amd.basis import prepare_basis
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.packing import pack
    specs = M.batch_water_clusters(256, seed=256, base_seed=1000, nmax=8, box=8.0)
    tabs = [pack(*prepare_basis(s, "6-31g*", cart=True)) for s in specs]
    plan = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
    assert plan.nmol == 256 and int(plan.mol_n.sum()) == 22458
    d1 = plan.pack([torch.as_tensor(systems.random_density(int(n), seed=100 + m)[0], device="cuda")
                    for m, n in enumerate(plan.mol_n)])
    d2 = plan.pack([torch.as_tensor(systems.random_density(int(n), seed=900 + m)[0], device="cuda")
                    for m, n in enumerate(plan.mol_n)])
    a = plan.jk(d1[None], symmetric=True)
    b = plan.jk(d2[None], symmetric=True)
    ab = plan.jk((d1 + 2.0 * d2)[None], symmetric=True)
    scale = float(ab.abs().max())
    assert float((ab - (a + 2.0 * b)).abs().max()) < 1e-12 * max(scale, 1.0)
    for j, k in zip(plan.unpack(a[0, 0]), plan.unpack(a[1, 0])):
        assert float((j - j.T).abs().max()) < 1e-12 and float((k - k.T).abs().max()) < 1e-12
    plan.incore(True)
    a_in = plan.jk(d1[None], symmetric=True)
    assert float((a_in - a).abs().max()) < 1e-12
    plan.incore(False)
