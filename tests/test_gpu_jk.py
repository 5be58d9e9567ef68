"""GPU parity of the HIP J/K engine against the oracle (and through it the
reference's libcint arithmetic), in the style of the reference's own tests:
  tests/integrals/test_native_direct_parallel.py:53-102 (J/K vs full-ERI contraction,
  nonsymmetric batched densities, Cartesian and spherical, screening linearity,
  huge cutoff -> zero) and :132-165 (operator selection).
Tolerances: 1e-10 Eh absolute per J/K element (the FP64 parity bar), 1e-12
absolute on individual ERI blocks.
"""
import itertools

import numpy as np
import pytest

from tests import systems

pytestmark = pytest.mark.gpu


def _cuda(x):
    import torch
    return torch.as_tensor(np.asarray(x), dtype=torch.float64, device="cuda")


def _plan(atm, bas, env, cart, cutoff=0.0):
    from gradscf_amd.jk import DirectJKPlan
    return DirectJKPlan(atm, bas, env, cart=cart, cutoff=cutoff)


def test_eri_blocks_every_class(oracle_mod):
    """Every canonical (l_i l_j|l_k l_l) class, general contraction included, vs the oracle."""
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=True)
    plan = _plan(atm, bas, env, True)
    ls = bas[:, 1]
    seen, quartets = set(), []
    rng = np.random.default_rng(3)
    for _ in range(20000):
        q = tuple(int(v) for v in rng.integers(0, len(bas), 4))
        key = tuple(int(ls[s]) for s in q)
        if key not in seen:
            seen.add(key)
            quartets.append(q)
    assert len(seen) == 81  # every ordered combination of s/p/d on four centres
    got = plan.eri_blocks(quartets)
    worst = 0.0
    for q, g in zip(quartets, got):
        ref = oracle_mod.eri_block(q, atm, bas, env, True)
        worst = max(worst, float(np.max(np.abs(g - ref))))
    assert worst < 1e-12, worst


@pytest.mark.parametrize("basis,cart,gc", [("6-31g", True, 0), ("cc-pvdz", True, 0), ("cc-pvdz", False, 0),
                                           ("cc-pvdz", True, 1), ("cc-pvdz", False, 1)])
def test_direct_jk_matches_oracle(oracle_mod, basis, cart, gc):
    """gc = 1: the O 1s/2s general contraction evaluated natively (one
    primitive quartet for both columns) instead of as two segments."""
    from gradscf_amd import autotune as at
    from gradscf_amd.jk import direct_jk
    _, _, atm, bas, env = systems.tables(systems.water(), basis, cart=cart)
    n = oracle_mod.nao(bas, cart)
    d = systems.random_density(n, nb=2, seed=1, symmetric=False)
    ref = oracle_mod.direct_jk(atm, bas, env, d, cart=cart, cutoff=0.0, workers=1)
    before = at.get_options()
    try:
        at.set_options(general_contraction=gc)
        j, k = direct_jk(atm, bas, env, _cuda(d), cart=cart, cutoff=0.0)
    finally:
        at.set_options(**before)
    assert np.max(np.abs(j.cpu().numpy() - ref[0])) < 1e-10
    assert np.max(np.abs(k.cpu().numpy() - ref[1])) < 1e-10


@pytest.mark.parametrize("cart", [True, False])
def test_direct_jk_screened_dimer(oracle_mod, cart):
    from gradscf_amd.jk import direct_jk
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=cart)
    n = oracle_mod.nao(bas, cart)
    d = systems.random_density(n, nb=1, seed=2, symmetric=True)[0]
    ref = oracle_mod.direct_jk(atm, bas, env, d, cart=cart, cutoff=1e-12, workers=4)
    j, k = direct_jk(atm, bas, env, _cuda(d), cart=cart, cutoff=1e-12)
    assert np.max(np.abs(j.cpu().numpy() - ref[0, 0])) < 1e-10
    assert np.max(np.abs(k.cpu().numpy() - ref[1, 0])) < 1e-10


def test_operator_selection_and_linearity(oracle_mod):
    import torch
    from gradscf_amd.jk import direct_jk
    _, _, atm, bas, env = systems.tables(systems.water(), "6-31g*", cart=True)
    n = oracle_mod.nao(bas, True)
    d1 = _cuda(systems.random_density(n, seed=4)[0])
    d2 = _cuda(systems.random_density(n, seed=5)[0])
    j1, k1 = direct_jk(atm, bas, env, d1, cart=True)
    j2, k2 = direct_jk(atm, bas, env, d2, cart=True)
    j3, k3 = direct_jk(atm, bas, env, 2.0 * d1 - 0.5 * d2, cart=True)
    assert torch.max(torch.abs(j3 - (2 * j1 - 0.5 * j2))) < 1e-10
    assert torch.max(torch.abs(k3 - (2 * k1 - 0.5 * k2))) < 1e-10
    # FP64 atomics sum in a run-dependent order: equal to rounding, not bitwise
    jo, ko = direct_jk(atm, bas, env, d1, cart=True, with_k=False)
    assert torch.max(torch.abs(jo - j1)) < 1e-12 and torch.count_nonzero(ko) == 0
    jo, ko = direct_jk(atm, bas, env, d1, cart=True, with_j=False)
    assert torch.max(torch.abs(ko - k1)) < 1e-12 and torch.count_nonzero(jo) == 0
    jz, kz = direct_jk(atm, bas, env, d1, cart=True, with_j=False, with_k=False)
    assert torch.count_nonzero(jz) == 0 and torch.count_nonzero(kz) == 0


def test_huge_cutoff_gives_zero():
    import torch
    from gradscf_amd.jk import direct_jk
    _, _, atm, bas, env = systems.tables(systems.water(), "6-31g", cart=True)
    d = _cuda(systems.random_density(13, seed=6)[0])
    j, k = direct_jk(atm, bas, env, d, cart=True, cutoff=1e30)
    assert torch.count_nonzero(j) == 0 and torch.count_nonzero(k) == 0


def test_complex_density(oracle_mod):
    import torch
    from gradscf_amd.jk import direct_jk
    _, _, atm, bas, env = systems.tables(systems.water(), "6-31g", cart=True)
    re = systems.random_density(13, seed=8, symmetric=False)[0]
    im = systems.random_density(13, seed=9, symmetric=False)[0]
    dc = torch.as_tensor(re + 1j * im, dtype=torch.complex128, device="cuda")
    j, k = direct_jk(atm, bas, env, dc, cart=True)
    ref_r = oracle_mod.direct_jk(atm, bas, env, re, cart=True)
    ref_i = oracle_mod.direct_jk(atm, bas, env, im, cart=True)
    assert np.max(np.abs(j.cpu().numpy() - (ref_r[0, 0] + 1j * ref_i[0, 0]))) < 1e-10
    assert np.max(np.abs(k.cpu().numpy() - (ref_r[1, 0] + 1j * ref_i[1, 0]))) < 1e-10


def test_invalid_inputs_rejected_before_launch():
    import torch
    from gradscf_amd.jk import direct_jk
    _, _, atm, bas, env = systems.tables(systems.water(), "6-31g", cart=True)
    d = _cuda(systems.random_density(13, seed=6)[0])
    with pytest.raises(ValueError, match="workers"):
        direct_jk(atm, bas, env, d, cart=True, workers=0)
    bad = d.clone()
    bad[0, 0] = float("nan")
    with pytest.raises(ValueError, match="Nonfinite density"):
        direct_jk(atm, bas, env, bad, cart=True)
    bas2 = bas.copy()
    bas2[0, 2] = 0
    with pytest.raises(ValueError, match="Invalid or unsupported libcint basis table"):
        direct_jk(atm, bas2, env, d, cart=True)
    env2 = env.copy()
    env2[int(bas[0, 5])] = -1.0
    with pytest.raises(ValueError, match="exponents must be positive"):
        direct_jk(atm, bas, env2, d, cart=True)
    with pytest.raises(ValueError, match="shape"):
        direct_jk(atm, bas, env, torch.zeros((12, 12), dtype=torch.float64, device="cuda"), cart=True)


def test_repeat_agrees_to_rounding(oracle_mod):
    """Same plan, same density, default FP64 atomics: repeated builds agree to
    the rounding of a reordered FP64 sum (the deterministic mode below makes
    them bitwise equal)."""
    from gradscf_amd.jk import direct_jk
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "6-31g*", cart=True)
    n = oracle_mod.nao(bas, True)
    d = _cuda(systems.random_density(n, seed=11)[0])
    a = direct_jk(atm, bas, env, d, cart=True, cutoff=1e-12)
    b = direct_jk(atm, bas, env, d, cart=True, cutoff=1e-12)
    for x, y in zip(a, b):
        assert float((x - y).abs().max()) < 1e-13


@pytest.mark.parametrize("cart", [True, False])
def test_schwarz_bounds_match_oracle(oracle_mod, cart):
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=cart)
    plan = _plan(atm, bas, env, cart, cutoff=1e-12)
    q = plan.schwarz()
    ref = oracle_mod.schwarz(atm, bas, env, cart)
    low = np.tril_indices(len(bas))
    assert np.max(np.abs(q[low] - ref[low]) / np.maximum(ref[low], 1e-300)) < 1e-12


@pytest.mark.parametrize("cart", [True, False])
def test_int1e_matches_oracle(oracle_mod, cart):
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=cart)
    plan = _plan(atm, bas, env, cart)
    got = plan.int1e()
    ref = oracle_mod.int1e(atm, bas, env, cart)
    for g, r in zip(got, ref):
        assert np.max(np.abs(g.cpu().numpy() - r)) < 1e-11


@pytest.mark.parametrize("basis,energy", [("6-31g*", -76.010492818245), ("sto-3g", -74.963082931333)])
def test_rhf_published_energies(basis, energy):
    """Reference-published RHF energies (examples/geometry/water_correlated_forces.py:84-87)."""
    from gradscf_amd.scf import RHF
    r = RHF("O 0 0 0; H 0 1.43 1.11; H 0 -1.43 1.11", basis, unit="Bohr", cart=True).run()
    assert r.converged
    assert abs(r.energy - energy) < 1e-9


def test_rhf_water_631g_matches_oracle(oracle_mod):
    """Water / 6-31G (O at origin, H at (0, +-0.757, 0.587) A): GPU SCF vs oracle SCF."""
    from oracle import scf as oscf
    from gradscf_amd.scf import RHF
    spec = systems.water()
    _, _, atm, bas, env = systems.tables(spec, "6-31g", cart=True)
    s, t, v = oracle_mod.int1e(atm, bas, env, True)

    def jk(d):
        o = oracle_mod.direct_jk(atm, bas, env, d, cart=True)
        return o[0, 0], o[1, 0]

    ref = oscf.rhf(t + v, s, spec.nelectron, spec.nuclear_repulsion(), jk)
    r = RHF(spec, "6-31g", cart=True).run()
    assert r.converged and ref["converged"]
    assert abs(r.energy - ref["energy"]) < 1e-9


def test_gpu_matches_golden_fixtures():
    """GPU J/K against fixtures computed on the reference libcint (tests/golden/make_golden.py)."""
    import json
    from pathlib import Path
    from gradscf_amd import molecule as M
    from gradscf_amd.jk import direct_jk
    from gradscf_amd.scf import RHF
    g = Path(__file__).resolve().parent / "golden"
    f = np.load(g / "water_631g_jk.npz")
    _, _, atm, bas, env = systems.tables(M.water_monomer(), "6-31g", cart=True)
    j, k = direct_jk(atm, bas, env, _cuda(f["density"]), cart=True)
    assert np.max(np.abs(j.cpu().numpy() - f["J"])) < 1e-10
    assert np.max(np.abs(k.cpu().numpy() - f["K"])) < 1e-10
    f = np.load(g / "dimer_ccpvdz_sph_jk.npz")
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=False)
    j, k = direct_jk(atm, bas, env, _cuda(f["density"]), cart=False, cutoff=1e-12)
    assert np.max(np.abs(j.cpu().numpy() - f["J"])) < 1e-10
    assert np.max(np.abs(k.cpu().numpy() - f["K"])) < 1e-10
    f = np.load(g / "water16_ccpvdz_sph_sample.npz")
    _, _, atm, bas, env = systems.tables(M.water_cluster(16, seed=16, box=12.0), "cc-pvdz", cart=False)
    c = np.random.default_rng(16).normal(size=(384, 80)) / np.sqrt(384)
    d = c @ c.T
    j, k = direct_jk(atm, bas, env, _cuda(d), cart=False, cutoff=1e-12)
    j, k = j.cpu().numpy(), k.cpu().numpy()
    idx = f["idx"]
    assert np.max(np.abs(j[idx[:, 0], idx[:, 1]] - f["J"])) < 1e-10
    assert np.max(np.abs(k[idx[:, 0], idx[:, 1]] - f["K"])) < 1e-10
    assert abs(np.sum(d * j) - f["trDJ"]) < 1e-9 * abs(f["trDJ"])
    assert abs(np.sum(d * k) - f["trDK"]) < 1e-9 * abs(f["trDK"])
    e = json.loads((g / "scf_energies.json").read_text())["water_631g_cart"]["energy"]
    assert abs(RHF(M.water_monomer(), "6-31g", cart=True).kernel() - e) < 1e-9


def test_mixed_precision_jk_close_to_fp64():
    """FP32 quartets below the threshold: J/K stay close to FP64; threshold 0 reproduces FP64."""
    import torch
    from gradscf_amd import molecule as M
    from gradscf_amd.jk import DirectJKPlan
    _, _, atm, bas, env = systems.tables(M.water_cluster(4, seed=3, box=6.0), "6-31g*", cart=True)
    plan = DirectJKPlan(atm, bas, env, cart=True, cutoff=1e-12)
    n = plan.nao
    c = torch.as_tensor(np.random.default_rng(5).normal(size=(n, 20)) / np.sqrt(n), device="cuda")
    d = (c @ c.T)[None].contiguous()
    ref = plan.jk(d, symmetric=True)
    same = plan.jk(d, symmetric=True, mixed_threshold=0.0)
    assert float((same - ref).abs().max()) < 1e-12
    mixed = plan.jk(d, symmetric=True, mixed_threshold=1e-7)
    err = float((mixed - ref).abs().max())
    assert err < 1e-6, err
    loose = plan.jk(d, symmetric=True, mixed_threshold=1e30)   # every quartet in FP32
    err_all = float((loose - ref).abs().max())
    assert err_all < 1e-3 and err_all > err


def test_mixed_precision_scf_energy_within_micro_hartree():
    from gradscf_amd import molecule as M
    from gradscf_amd.scf import RHF
    spec = M.water_cluster(4, seed=3, box=6.0)
    e64 = RHF(spec, "6-31g*", cart=True, direct_scf_tol=1e-12).kernel()
    emix = RHF(spec, "6-31g*", cart=True, direct_scf_tol=1e-12, mixed_threshold=1e-7).kernel()
    assert abs(emix - e64) < 1e-6, (emix, e64)


@pytest.mark.parametrize("basis,cart", [("6-31g*", True), ("cc-pvdz", False)])
def test_geometry_vjp_matches_oracle(oracle_mod, basis, cart):
    """Coordinate VJP of J/K vs the oracle's s1 loop over libcint-convention int2e_ip1
    (csrc/direct_jk_geometry.cc:68-84), nonsymmetric cotangents."""
    from gradscf_amd.jk import DirectJKPlan
    _, _, atm, bas, env = systems.tables(systems.water(), basis, cart=cart)
    n = oracle_mod.nao(bas, cart)
    rng = np.random.default_rng(21)
    d = systems.random_density(n, seed=22)[0]
    cj = rng.normal(size=(n, n))
    ck = rng.normal(size=(n, n))
    plan = DirectJKPlan(atm, bas, env, cart=cart, cutoff=0.0)
    got = plan.jk_geometry_vjp(_cuda(d), _cuda(cj), _cuda(ck)).cpu().numpy()
    ref = oracle_mod.jk_geometry_vjp(atm, bas, env, d, cj, ck, cart=cart)
    assert np.max(np.abs(got - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(got.sum(axis=0))) < 1e-9  # translation sum rule


def test_geometry_vjp_finite_difference():
    """E(R) = tr(D J)/2 - tr(D K)/4 at fixed D: VJP gradient vs central differences
    (step 1e-4 bohr) for a few coordinate slots."""
    import torch
    from gradscf_amd.jk import DirectJKPlan, direct_jk
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "6-31g*", cart=True)
    plan = DirectJKPlan(atm, bas, env, cart=True, cutoff=0.0)
    n = plan.nao
    c = np.random.default_rng(3).normal(size=(n, 10)) / np.sqrt(n)
    d = c @ c.T
    dt = _cuda(d)
    g = plan.jk_geometry_vjp(dt, 0.5 * dt, -0.25 * dt).cpu().numpy()

    def energy(env_):
        j, k = direct_jk(atm, bas, env_, dt, cart=True)
        return float(0.5 * torch.sum(dt * j) - 0.25 * torch.sum(dt * k))

    h = 1e-4
    for row in (6, 9, 14):
        for x in range(3):
            e1, e2 = env.copy(), env.copy()
            e1[atm[row, 1] + x] += h
            e2[atm[row, 1] + x] -= h
            fd = (energy(e1) - energy(e2)) / (2 * h)
            assert abs(fd - g[row, x]) < 2e-7, (row, x, fd, g[row, x])


def test_screened_geometry_vjp_rejected():
    from gradscf_amd.jk import DirectJKPlan
    _, _, atm, bas, env = systems.tables(systems.water(), "6-31g", cart=True)
    plan = DirectJKPlan(atm, bas, env, cart=True, cutoff=1e-12)
    d = _cuda(np.eye(13))
    with pytest.raises(ValueError, match="cutoff=0"):
        plan.jk_geometry_vjp(d, d, None)


@pytest.mark.parametrize("basis,cart,cutoff", [("6-31g*", True, 1e-12), ("cc-pvdz", False, 0.0)])
def test_workspace_forecast_bounds_plan_memory(basis, cart, cutoff):
    from gradscf_amd.basis import prepare_basis
    from gradscf_amd.jk import DirectJKPlan, NativeDirectBasis, make_plan, native_jk_workspace_bytes
    from gradscf_amd.packing import pack
    topo, par = prepare_basis(systems.water_dimer(), basis, cart=cart)
    atm, bas, env = pack(topo, par)
    plan = DirectJKPlan(atm, bas, env, cart=cart, cutoff=cutoff)
    d = _cuda(systems.random_density(plan.nao, nb=2, seed=5, symmetric=False))
    plan.jk(d)
    import ctypes
    from gradscf_amd import _native
    info = _native.PlanInfo()
    _native.check(_native.lib().gscf_plan_get_info(plan._handle, ctypes.byref(info)))
    forecast = native_jk_workspace_bytes(NativeDirectBasis(make_plan(topo), par), 1, cutoff, nb=2)
    assert 0 < info.device_bytes <= forecast


@pytest.mark.parametrize("sched", [dict(rys_lds=1), dict(chunk_items=1024, bras_max=4, target_prim=12),
                                   dict(rys_lds=1, chunk_items=8192, bras_max=32, target_prim=96),
                                   dict(ket_lds=1), dict(ket_lds=1, threads=512, rys_lds=1, chunk_items=1024),
                                   dict(threads=512, bras_max=1)])
def test_schedules_do_not_change_results(sched):
    """Every tuning-table schedule (Rys table in LDS, chunk sizes, primitive
    split) gives the default schedule's J/K up to FP64 summation order."""
    from gradscf_amd import autotune as at
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=False)
    at.reset()
    ref_plan = _plan(atm, bas, env, False, 1e-12)
    d = _cuda(systems.random_density(ref_plan.nao, nb=2, seed=9, symmetric=False))
    ref = ref_plan.jk(d)
    ref_m = ref_plan.jk(d[:1] * 1e-3, symmetric=False, mixed_threshold=1e-7)
    try:
        at.set_schedule(at.Schedule(**sched))
        plan = _plan(atm, bas, env, False, 1e-12)
        assert float((plan.jk(d) - ref).abs().max()) < 1e-12
        got_m = plan.jk(d[:1] * 1e-3, symmetric=False, mixed_threshold=1e-7)
        # per-wavefront precision choice: FP32 rounding of sub-threshold quartets
        assert float((got_m - ref_m).abs().max()) < 1e-11
    finally:
        at.reset()


@pytest.mark.parametrize("with_j,with_k", [(True, True), (True, False), (False, True)])
def test_many_nonsymmetric_densities_match_single_builds(with_j, with_k):
    """(H2O)10 cc-pVDZ (250 Cartesian AO) with native general contraction: two
    nonsymmetric densities in one build exceed the LDS of the ket-general
    classes, which then run one density per launch; the result must equal two
    single-density builds (operator selection included)."""
    import torch
    from gradscf_amd import autotune as at
    from gradscf_amd import molecule as M
    _, _, atm, bas, env = systems.tables(M.water_cluster(10, seed=10, box=9.0), "cc-pvdz", cart=True)
    before = at.get_options()
    try:
        at.set_options(general_contraction=1)
        plan = _plan(atm, bas, env, True, 1e-12)
    finally:
        at.set_options(**before)
    d = _cuda(systems.random_density(plan.nao, nb=2, seed=12, symmetric=False))
    both = plan.jk(d, with_j=with_j, with_k=with_k)
    for b in range(2):
        one = plan.jk(d[b:b + 1], with_j=with_j, with_k=with_k)
        assert float((both[:, b] - one[:, 0]).abs().max()) < 1e-12


@pytest.mark.parametrize("pm", [0, 2, 3])
def test_partial_sum_flushes_do_not_change_results(pm):
    """(H2O)4 cc-pVDZ spherical, nonsymmetric densities: exchange rows and
    ket-side J windows handed to the reduction pass as partial slabs (plain
    stores) instead of flushed with device-scope atomics, for J and K
    together and alone, evaluated and in-core, change J/K only by the order
    of FP64 additions."""
    import torch
    from gradscf_amd import autotune as at
    from gradscf_amd import molecule as M
    _, _, atm, bas, env = systems.tables(M.water_cluster(4, seed=4, box=7.0), "cc-pvdz", cart=False)
    ref_plan = _plan(atm, bas, env, False, 1e-12)
    d = _cuda(systems.random_density(ref_plan.nao, nb=2, seed=6, symmetric=False))
    refs = {(wj, wk): ref_plan.jk(d, with_j=wj, with_k=wk) for wj, wk in ((1, 1), (1, 0), (0, 1))}
    before = at.get_options()
    try:
        at.set_options(partial_sums=pm)
        plan = _plan(atm, bas, env, False, 1e-12)
        for incore in (False, True):
            if incore:
                plan.incore(True)
            for (wj, wk), ref in refs.items():
                got = plan.jk(d, with_j=wj, with_k=wk)
                assert float((got - ref).abs().max()) < 1e-11, (pm, incore, wj, wk)
    finally:
        at.set_options(**before)


def test_partial_sum_flushes_on_a_batched_plan():
    """Both flush kinds as partial slabs on a batched plan (molecules of
    different widths share the slabs and the per-stream row lists)."""
    import torch
    from gradscf_amd import autotune as at
    from gradscf_amd import molecule as M
    from gradscf_amd.jk import DirectJKPlan
    specs = [M.water_monomer(), M.water_cluster(3, seed=1003, box=8.0), M.water_cluster(2, seed=1002, box=8.0)]
    tabs = [systems.tables(sp, "6-31g*", cart=True)[2:] for sp in specs]
    before = at.get_options()
    outs = {}
    try:
        for pm in (0, 3):
            at.set_options(partial_sums=pm)
            plan = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
            dens = plan.pack([torch.as_tensor(systems.random_density(int(n), nb=2, seed=60 + m, symmetric=False),
                                              device="cuda") for m, n in enumerate(plan.mol_n)])
            outs[pm] = plan.jk(dens)
    finally:
        at.set_options(**before)
    assert float((outs[3] - outs[0]).abs().max()) < 1e-11


@pytest.mark.parametrize("gc,streams", [(1, 4), (1, 1), (0, 1)])
def test_engine_options_do_not_change_results(gc, streams):
    """(H2O)4 cc-pVDZ spherical: native general contraction instead of
    segments (the default, checked against the oracle elsewhere), and class
    launches on one stream instead of four, change J/K only by the order of
    FP64 additions (nonsymmetric density, J and K)."""
    import torch
    from gradscf_amd import autotune as at
    from gradscf_amd import molecule as M
    _, _, atm, bas, env = systems.tables(M.water_cluster(4, seed=4, box=7.0), "cc-pvdz", cart=False)
    ref_plan = _plan(atm, bas, env, False, 1e-12)
    d = _cuda(systems.random_density(ref_plan.nao, nb=1, seed=5, symmetric=False))
    ref = ref_plan.jk(d)
    before = at.get_options()
    try:
        at.set_options(general_contraction=gc, streams=streams)
        plan = _plan(atm, bas, env, False, 1e-12)
        got = plan.jk(d)
        got2 = plan.jk(d)
    finally:
        at.set_options(**before)
    assert float((got - ref).abs().max()) < 1e-11
    assert float((got2 - got).abs().max()) < 1e-11


@pytest.mark.parametrize("lds_limit,route", [(0, "all densities"), (8192, "per density"), (4096, "global atomics")])
def test_lds_routes_match_oracle(oracle_mod, lds_limit, route):
    """Water dimer cc-pVDZ (Cartesian, nonsymmetric densities): the class
    launches an LDS limit forces -- every density in one launch, one launch
    per density (the route of wide molecules), and the global-atomic kernels
    for classes whose exchange rows do not fit even for one density -- give
    the oracle's J/K, evaluated and in core, for J and K together and alone."""
    from gradscf_amd import autotune as at
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=True)
    n = oracle_mod.nao(bas, True)
    d = systems.random_density(n, nb=2, seed=71, symmetric=False)
    ref = oracle_mod.direct_jk(atm, bas, env, d, cart=True, cutoff=1e-12, workers=4)
    before = at.get_options()
    try:
        at.set_options(lds_limit=lds_limit)
        plan = _plan(atm, bas, env, True, 1e-12)
    finally:
        at.set_options(**before)
    assert (plan.info["classes_global_atomic"] > 0) == (route == "global atomics")
    if lds_limit:
        assert plan.info["lds_max"] == lds_limit
    dd = _cuda(d)
    for incore in (False, True):
        if incore:
            plan.incore(True)
        out = plan.jk(dd).cpu().numpy()
        assert np.max(np.abs(out - ref)) < 1e-10, (route, incore)
        j = plan.jk(dd, with_k=False).cpu().numpy()
        k = plan.jk(dd, with_j=False).cpu().numpy()
        assert np.max(np.abs(j[0] - ref[0])) < 1e-10 and np.max(np.abs(k[0] - ref[1])) < 1e-10


def test_nonsymmetric_density_on_the_bench_workload(oracle_mod):
    """(H2O)16 cc-pVDZ spherical (400 Cartesian AOs), one nonsymmetric density:
    the dd classes' exchange rows for D and D^T exceed 64 KB together and run
    one density per launch; J/K against the reference libcint path."""
    from gradscf_amd import molecule as M
    _, _, atm, bas, env = systems.tables(M.water_cluster(16, seed=16, box=12.0), "cc-pvdz", cart=False)
    plan = _plan(atm, bas, env, False, 1e-12)
    assert plan.info["classes_global_atomic"] == 0
    d = systems.random_density(plan.nao, nb=1, seed=72, symmetric=False) * 0.05
    with oracle_mod.backend("libcint" if oracle_mod.have_ref() else "oracle"):
        ref = oracle_mod.direct_jk(atm, bas, env, d, cart=False, cutoff=1e-12, workers=16)
    out = plan.jk(_cuda(d)).cpu().numpy()
    assert np.max(np.abs(out - ref)) < 1e-10


def test_wide_molecule_one_density_slot_above_64kb(oracle_mod):
    """(H2O)36 6-31G* (684 Cartesian AOs): one bra slot of the dd classes needs
    more than 64 KB of LDS (12 exchange rows x 684 columns), so those classes
    run one density per launch with the larger LDS the device allows instead
    of the global-atomic kernels."""
    from gradscf_amd import molecule as M
    _, _, atm, bas, env = systems.tables(M.water_cluster(36, seed=36, box=16.0), "6-31g*", cart=True)
    plan = _plan(atm, bas, env, True, 1e-12)
    assert plan.info["nc_max"] == 684
    assert plan.info["lds_max"] > 65536 and plan.info["classes_global_atomic"] == 0
    c = np.random.default_rng(73).normal(size=(684, 180)) / np.sqrt(684)
    d = (c @ c.T)[None]
    with oracle_mod.backend("libcint" if oracle_mod.have_ref() else "oracle"):
        ref = oracle_mod.direct_jk(atm, bas, env, d, cart=True, cutoff=1e-12, workers=16)
    out = plan.jk(_cuda(d)).cpu().numpy()
    assert np.max(np.abs(out - ref)) < 1e-10


@pytest.mark.parametrize("cart,lds_limit", [(True, 0), (False, 0), (True, 8192), (True, 4096)])
def test_deterministic_mode_is_bitwise_reproducible(oracle_mod, cart, lds_limit):
    """Engine option deterministic=1 (integer accumulation of scaled values):
    repeated builds, builds on one launch stream instead of four and in-core
    builds of the same plan are bitwise equal, on every class route (all
    densities in one launch, one per density, global atomics), and equal the
    oracle to 1e-10 (the reference is reproducible at a fixed worker count,
    tests/integrals/test_native_direct_parallel.py:63-70)."""
    import torch
    from gradscf_amd import autotune as at
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=cart)
    n = oracle_mod.nao(bas, cart)
    d = systems.random_density(n, nb=2, seed=802, symmetric=False)
    ref = oracle_mod.direct_jk(atm, bas, env, d, cart=cart, cutoff=1e-12, workers=4)
    before = at.get_options()
    outs = []
    try:
        for streams in (4, 1):
            at.set_options(deterministic=1, lds_limit=lds_limit, streams=streams)
            plan = _plan(atm, bas, env, cart, 1e-12)
            dd = _cuda(d)
            outs += [plan.jk(dd) for _ in range(3)]
            if streams == 4:
                plan.incore(True)
                outs += [plan.jk(dd) for _ in range(2)]
                plan.incore(False)
                with pytest.raises(ValueError, match="deterministic"):
                    plan.jk(dd, mixed_threshold=1e-7)
    finally:
        at.set_options(**before)
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
    assert np.max(np.abs(outs[0].cpu().numpy() - ref)) < 1e-10


def test_deterministic_mode_batched_plan():
    """A batched plan (every molecule its own fixed-point scale): bitwise
    repeatable, and equal to the FP64-atomic build to 1e-11."""
    import torch
    from gradscf_amd import autotune as at
    from gradscf_amd import molecule as M
    from gradscf_amd.jk import DirectJKPlan
    specs = [M.water_monomer(), M.water_cluster(3, seed=1003, box=8.0), M.water_cluster(2, seed=1002, box=8.0)]
    tabs = [systems.tables(sp, "6-31g*", cart=True)[2:] for sp in specs]
    plan64 = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
    dens = plan64.pack([torch.as_tensor(systems.random_density(int(n), nb=1, seed=70 + m), device="cuda")
                        for m, n in enumerate(plan64.mol_n)])
    ref = plan64.jk(dens)
    before = at.get_options()
    try:
        at.set_options(deterministic=1)
        plan = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
    finally:
        at.set_options(**before)
    a, b = plan.jk(dens), plan.jk(dens)
    assert torch.equal(a, b)
    assert float((a - ref).abs().max()) < 1e-11
