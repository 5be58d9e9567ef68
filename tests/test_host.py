"""Host-side checks that need no GPU: basis parsing and packing conventions, the
workload geometry generators, the Rys interpolation table, and the C ABI
library (loads, exports every symbol include/gradscf_jk.h declares)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from gradscf_amd import molecule as M
from gradscf_amd.basis import normalized_shell_coefficients, prepare_basis
from gradscf_amd.packing import ao_offsets, pack
from tests import systems

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("spec,basis,cart,nshell,nao", [
    (M.water_monomer(), "6-31g", True, 9, 13),
    (M.water_cluster(16, seed=16, box=12.0), "cc-pvdz", False, 176, 384),
    (M.water_cluster(16, seed=16, box=12.0), "cc-pvdz", True, 176, 400),
    (M.linear_alkane(20), "6-31g*", True, 204, 384),
    (M.water_cluster(32, seed=32, box=15.0), "6-31g*", True, 320, 608),
])
def test_workload_sizes(spec, basis, cart, nshell, nao):
    """Shell and AO counts of the benchmark systems (basis compositions of the
    reference's PySCF snapshot: 6-31G O/C S6 S3 S1 P3 P1, H S3 S1; 6-31G* adds a
    Cartesian d on O/C; cc-pVDZ O S[8,2c] S1 P3 P1 D1, H S3 S1 P1)."""
    topo, _ = prepare_basis(spec, basis, cart=cart)
    assert topo.nshell == nshell
    assert topo.nao() == nao


def test_shell_order_within_atom():
    topo, par = prepare_basis(M.water_monomer(), "cc-pvdz")
    assert topo.angular_momenta[:5] == (0, 0, 1, 1, 2)
    assert topo.primitive_counts[:5] == (8, 1, 3, 1, 1)
    assert topo.contraction_counts[:5] == (2, 1, 1, 1, 1)
    assert topo.angular_momenta[5:] == (0, 0, 1, 0, 0, 1)


def test_normalized_coefficients_unit_norm():
    """Contracted radial functions r^l sum c exp(-a r^2) have unit norm."""
    import math
    topo, par = prepare_basis(M.water_monomer(), "cc-pvdz")
    for l, a, c in zip(topo.angular_momenta, par.exponents, par.coefficients):
        cn = normalized_shell_coefficients(l, a, c)
        for col in range(cn.shape[1]):
            s = 0.0
            for p in range(len(a)):
                for q in range(len(a)):
                    s += cn[p, col] * cn[q, col] * math.gamma(l + 1.5) / (2 * (a[p] + a[q]) ** (l + 1.5))
            assert abs(s - 1.0) < 1e-13


def test_pack_layout():
    topo, par = prepare_basis(M.water_monomer(), "6-31g")
    atm, bas, env = pack(topo, par)
    assert atm.shape == (3 + 9, 6) and bas.shape == (9, 8)
    assert list(atm[:3, 0]) == [8, 1, 1] and np.all(atm[3:, 0] == 0) and np.all(atm[:, 2] == 1)
    assert list(atm[:, 1]) == [20 + 3 * i for i in range(12)]
    assert bas[0, 0] == 3 and bas[0, 5] == 20 + 36 and bas[0, 6] == bas[0, 5] + 6
    np.testing.assert_allclose(env[atm[4, 1]:atm[4, 1] + 3], par.coords_bohr[0] if hasattr(par, "coords_bohr")
                               else par.centers[1])
    assert int(ao_offsets(bas, True)[-1]) == 13


def test_geometries_match_fixtures():
    for name, spec in [("water", M.water_monomer()),
                       ("water16", M.water_cluster(16, seed=16, box=12.0)),
                       ("water32", M.water_cluster(32, seed=32, box=15.0)),
                       ("c20h42", M.linear_alkane(20))]:
        ref = M.from_xyz((GOLDEN / f"{name}.xyz").read_text())
        assert ref.symbols == spec.symbols
        np.testing.assert_allclose(spec.coords_bohr, ref.coords_bohr, atol=1e-9)


def test_water_cluster_rules():
    spec = M.water_cluster(16, seed=16, box=12.0)
    r = spec.coords_bohr / M.ANGSTROM_TO_BOHR
    o = r[0::3]
    assert np.all((o >= 0) & (o <= 12.0))
    d = np.linalg.norm(o[:, None] - o[None], axis=-1) + np.eye(16) * 100
    assert d.min() >= 2.6
    for w in range(16):
        oh1, oh2 = r[3 * w + 1] - r[3 * w], r[3 * w + 2] - r[3 * w]
        assert abs(np.linalg.norm(oh1) - 0.96) < 1e-12 and abs(np.linalg.norm(oh2) - 0.96) < 1e-12
        ang = np.degrees(np.arccos(oh1 @ oh2 / 0.96 ** 2))
        assert abs(ang - 104.5) < 1e-9


def test_alkane_geometry():
    spec = M.linear_alkane(20)
    r = spec.coords_bohr / M.ANGSTROM_TO_BOHR
    c, h = r[:20], r[20:]
    assert len(h) == 42
    assert np.allclose(np.linalg.norm(np.diff(c, axis=0), axis=1), 1.54)
    for hh in h:
        assert abs(np.min(np.linalg.norm(c - hh, axis=1)) - 1.09) < 1e-12
    v1, v2 = c[0] - c[1], c[2] - c[1]
    assert abs(np.degrees(np.arccos(v1 @ v2 / 1.54 ** 2)) - 109.4712206) < 1e-6


def test_rys_table_against_mpmath():
    """The generated interpolation table reproduces 40-digit roots/weights."""
    mpmath = pytest.importorskip("mpmath")
    import importlib.util
    path = ROOT / "gradscf_amd" / "csrc" / "gen_rys_tables.py"
    spec = importlib.util.spec_from_file_location("gen", path)
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    head = (ROOT / "gradscf_amd/csrc/_generated/rys_table.inc")
    body = (ROOT / "gradscf_amd/csrc/_generated/rys_table_body.inc")
    if not head.is_file():
        pytest.skip("engine not built")
    txt = head.read_text()
    off = [int(v) for v in re.search(r"rys_table_offset\[8\] = \{([^}]*)\}", txt).group(1).split(",")]
    tab = np.array([float.fromhex(v.strip()) for v in body.read_text().replace("\n", "").split(",") if v.strip()])
    rng = np.random.default_rng(1)
    for n in range(1, 8):
        for x in list(rng.uniform(0, 80, 6)) + [0.0, 79.999]:
            k = int(x / 2)
            y = x - (2 * k + 1)
            c = tab[off[n - 1] + k * 28 * n: off[n - 1] + (k + 1) * 28 * n].reshape(14, 2 * n)
            f = c[13].copy()
            for jj in range(12, -1, -1):   # Horner in y, as the device evaluates it
                f = f * y + c[jj]
            r, w = gen._rys_mp(n, float(x))
            assert np.max(np.abs(f[:n] - r) / np.array(r)) < 5e-15
            assert np.max(np.abs(f[n:] - w) / np.array(w)) < 5e-15


def test_boys_moment_quadrature_against_mpmath():
    """One- and two-point Rys quadrature from Taylor-interpolated Boys moments
    (the device's rys_roots_boys, restated operation by operation) against
    60-digit roots and weights, across the grid and at the switch to the
    large-x limit."""
    pytest.importorskip("mpmath")
    import importlib.util
    path = ROOT / "gradscf_amd" / "csrc" / "gen_rys_tables.py"
    spec = importlib.util.spec_from_file_location("gen", path)
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    body = ROOT / "gradscf_amd/csrc/_generated/boys_table_body.inc"
    head = ROOT / "gradscf_amd/csrc/_generated/rys_table.inc"
    if not body.is_file():
        pytest.skip("engine not built")
    tab = np.array([float.fromhex(v.strip()) for v in body.read_text().replace("\n", "").split(",") if v.strip()])
    txt = head.read_text()
    asym_r = [float.fromhex(v) for v in re.search(r"rys_asym_root_c\[28\] = \{([^}]*)\}", txt).group(1).split(",")]
    asym_w = [float.fromhex(v) for v in re.search(r"rys_asym_weight_c\[28\] = \{([^}]*)\}", txt).group(1).split(",")]
    H, K, NB, TMAX = gen.BOYS_H, gen.BOYS_K, gen.BOYS_NB, gen.BOYS_TMAX
    assert tab.size == (int(TMAX / H) + 1) * NB

    def boys(x, m_count):
        i = int(x * (1.0 / H) + 0.5)
        d = i * H - x
        f = tab[i * NB:(i + 1) * NB]
        dk = [0.0] + [d * (1.0 / k) for k in range(1, K)]
        out = []
        for m in range(m_count):
            r = f[m + K - 1]
            for k in range(K - 1, 0, -1):
                r = r * dk[k] + f[m + k - 1]
            out.append(r)
        return out

    def quad(n, x):
        if x >= TMAX:
            b = n * (n - 1) // 2
            return [asym_r[b + f] / x for f in range(n)], [asym_w[b + f] / np.sqrt(x) for f in range(n)]
        if n == 1:
            f0, f1 = boys(x, 2)
            return [f1 / f0], [f0]
        m0, m1, m2, m3 = boys(x, 4)
        i_d = 1.0 / (m0 * m2 - m1 * m1)
        a1 = (m1 * m2 - m0 * m3) * i_d
        a0 = (m1 * m3 - m2 * m2) * i_d
        s2 = 0.5 * (np.sqrt(a1 * a1 - 4.0 * a0) - a1)
        s1 = a0 / s2
        w1 = (s2 * m0 - m1) / (s2 - s1)
        return [s1, s2], [w1, m0 - w1]

    rng = np.random.default_rng(7)
    xs = list(rng.uniform(0, TMAX, 40)) + [0.0, 1e-13, H / 2 - 1e-12, H / 2 + 1e-12, 25.0, TMAX - 1e-9, TMAX, 60.0]
    for n in (1, 2):
        for x in xs:
            s, w = quad(n, float(x))
            r, ww = gen._rys_mp(n, float(x))
            assert max(abs(a - b) / b for a, b in zip(s, r)) < 5e-15, (n, x)
            assert max(abs(a - b) / b for a, b in zip(w, ww)) < 5e-15, (n, x)


def test_eigh_placement_and_host_path():
    """Small single eigen-problems go to host LAPACK, batches and large ones
    stay on the GPU; the host path returns the same decomposition as torch's,
    and the padded form of small batches the same eigenpairs."""
    import torch
    from gradscf_amd import linalg
    assert linalg.on_host((13, 13)) and linalg.on_host((384, 384)) and not linalg.on_host((608, 608))
    assert not linalg.on_host((32, 57, 57)) and not linalg.on_host((32, 76, 76))
    assert linalg.on_host((1, 384, 384))
    a = torch.randn(6, 6, dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    a = a + a.T
    e, v = linalg.eigh(a)
    assert torch.allclose(v @ torch.diag(e) @ v.T, a, atol=1e-12)
    b = torch.randn(5, 19, 19, dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    b = b + b.transpose(1, 2)
    ep, vp = linalg.eigh_padded(b, 65)
    e0, v0 = torch.linalg.eigh(b)
    assert torch.allclose(ep, e0, atol=1e-12)
    # eigenvectors up to sign: the reconstruction and the spectral projectors
    assert torch.allclose(vp @ torch.diag_embed(ep) @ vp.transpose(1, 2), b, atol=1e-11)
    assert torch.allclose(vp[..., :7] @ vp[..., :7].transpose(1, 2), v0[..., :7] @ v0[..., :7].transpose(1, 2),
                          atol=1e-10)


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "gradscf_jk.h").read_text()
    declared = set(re.findall(r"\b(gscf_[a-z0-9_]+)\s*\(", header))
    assert declared
    lib_path = ROOT / "gradscf_amd" / "lib" / "libgradscf_jk.so"
    if not lib_path.is_file():
        pytest.skip("engine not built")
    lib = ctypes.CDLL(str(lib_path), mode=ctypes.RTLD_LOCAL)
    for name in sorted(declared):
        assert hasattr(lib, name), name
    from gradscf_amd import _native
    assert set(_native.EXPORTS) == declared


def test_c_abi_rejects_bad_tables_without_gpu():
    """Host validation runs before any device call, so it works on a CPU-only host."""
    from gradscf_amd import _native
    lib_path = ROOT / "gradscf_amd" / "lib" / "libgradscf_jk.so"
    if not lib_path.is_file():
        pytest.skip("engine not built")
    L = _native.lib()
    _, _, atm, bas, env = systems.tables(M.water_monomer(), "6-31g")
    bad = bas.copy()
    bad[0, 1] = 3  # f shell: outside the engine's s/p/d support
    env_f = np.ascontiguousarray(env)
    h = ctypes.c_void_p()
    rc = L.gscf_plan_create(atm.ctypes.data, len(atm), bad.ctypes.data, len(bad), env_f.ctypes.data, len(env_f),
                            1, ctypes.c_double(0.0), None, ctypes.byref(h))
    assert rc == _native.GSCF_INVALID_ARGUMENT
    assert b"l <= 2" in L.gscf_last_error()
    env2 = env_f.copy()
    env2[8] = 0.3
    rc = L.gscf_plan_create(atm.ctypes.data, len(atm), bas.ctypes.data, len(bas), env2.ctypes.data, len(env2),
                            1, ctypes.c_double(0.0), None, ctypes.byref(h))
    assert rc == _native.GSCF_INVALID_ARGUMENT
    # the one-shot drop-in reports every failure as Internal, like the reference's
    # GradSCFDirectJK (direct_jk.cc:83,93), with the reference's messages
    def one_shot(b=bas, e=env_f, n=13, workers=1, with_j=1, with_k=1):
        return L.gscf_direct_jk(atm.ctypes.data, len(atm), b.ctypes.data, len(b), e.ctypes.data, len(e),
                                None, 1, n, None, 1, ctypes.c_double(0.0), workers, with_j, with_k, None)

    for kwargs, msg in (({"workers": 0}, b"workers must be positive"),
                        ({"with_j": 0, "with_k": 0}, b"at least one selected operator"),
                        ({"with_j": 2}, b"at least one selected operator"),
                        ({"n": 12}, b"shape mismatch"),
                        ({"b": bad}, b"l <= 2"),
                        ({"e": env2}, b"Range-separated")):
        assert one_shot(**kwargs) == _native.GSCF_INTERNAL, kwargs
        assert msg in L.gscf_last_error(), (kwargs, L.gscf_last_error())
    for blocks, iters in ((0, 10), (4, 0)):
        assert L.gscf_fma_probe(blocks, iters, ctypes.c_void_p(8), None) == _native.GSCF_INVALID_ARGUMENT
    assert L.gscf_fma_probe(4, 10, None, None) == _native.GSCF_INVALID_ARGUMENT


def test_concat_tables_preserves_each_molecule():
    """Batched tables: every molecule's shells read the same exponents,
    coefficients and centre as in its own tables, and sit on its own atoms."""
    from gradscf_amd.jk import concat_tables
    specs = [M.water_monomer(), M.water_cluster(2, seed=1002, box=8.0), M.water_cluster(1, seed=1001, box=8.0)]
    tabs = [systems.tables(s, "6-31g*")[2:] for s in specs]
    atm, bas, env, mol_atm, mol_bas = concat_tables(tabs)
    assert list(np.diff(mol_atm)) == [len(t[0]) for t in tabs]
    assert list(np.diff(mol_bas)) == [len(t[1]) for t in tabs] and mol_bas[-1] == len(bas)
    for m, (a0, b0, e0) in enumerate(tabs):
        for i, row in enumerate(b0):
            r = bas[mol_bas[m] + i]
            assert mol_atm[m] <= r[0] < mol_atm[m + 1]
            np.testing.assert_array_equal(env[r[5]:r[5] + r[2]], e0[row[5]:row[5] + row[2]])
            np.testing.assert_array_equal(env[r[6]:r[6] + r[2] * r[3]], e0[row[6]:row[6] + row[2] * row[3]])
            ra = atm[r[0]]
            np.testing.assert_array_equal(env[ra[1]:ra[1] + 3], e0[a0[row[0]][1]:a0[row[0]][1] + 3])


def test_c_abi_rejects_bad_batch_offsets_without_gpu():
    from gradscf_amd import _native
    from gradscf_amd.jk import concat_tables
    if not (ROOT / "gradscf_amd" / "lib" / "libgradscf_jk.so").is_file():
        pytest.skip("engine not built")
    L = _native.lib()
    tabs = [systems.tables(M.water_monomer(), "6-31g")[2:], systems.tables(M.water_monomer(), "sto-3g")[2:]]
    atm, bas, env, mol_atm, mol_bas = concat_tables(tabs)
    h = ctypes.c_void_p()

    def create(a, b, ma, mb, nmol=2):
        return L.gscf_plan_create_batch(a.ctypes.data, len(a), b.ctypes.data, len(b), env.ctypes.data, len(env),
                                        nmol, ma.ctypes.data, mb.ctypes.data, 1, ctypes.c_double(0.0), None,
                                        ctypes.byref(h))

    bad_end = mol_bas.copy()
    bad_end[-1] -= 1
    assert create(atm, bas, mol_atm, bad_end) == _native.GSCF_INVALID_ARGUMENT
    assert b"end at" in L.gscf_last_error()
    empty = mol_bas.copy()
    empty[1] = 0
    assert create(atm, bas, mol_atm, empty) == _native.GSCF_INVALID_ARGUMENT
    cross = bas.copy()
    cross[mol_bas[1], 0] = 0  # a shell of molecule 1 placed on an atom of molecule 0
    assert create(atm, cross, mol_atm, mol_bas) == _native.GSCF_INVALID_ARGUMENT
    assert b"basis table" in L.gscf_last_error()
    assert create(atm, bas, mol_atm, mol_bas, nmol=0) == _native.GSCF_INVALID_ARGUMENT


def test_workspace_forecast_without_gpu():
    """The memory forecast (the reference's native_jk_workspace_bytes) runs on
    a CPU-only host, validates like plan creation, grows with the density count
    and is None without a complete topology."""
    if not (ROOT / "gradscf_amd" / "lib" / "libgradscf_jk.so").is_file():
        pytest.skip("engine not built")
    from gradscf_amd.jk import NativeDirectBasis, make_plan, native_jk_workspace_bytes
    topo, par = prepare_basis(M.water_cluster(4, seed=4, box=6.0), "6-31g*", cart=True)
    prov = NativeDirectBasis(make_plan(topo), par)
    one = native_jk_workspace_bytes(prov, 1, 1e-12)
    two = native_jk_workspace_bytes(prov, 1, 1e-12, nb=2)
    four = native_jk_workspace_bytes(prov, 1, 1e-12, nb=4)
    n = topo.nao()
    assert one > 3 * 2 * n * n * 8
    # per extra density two more workspace slots (D and D^T for nonsymmetric K),
    # each with Cartesian D, Jp, Kp, the ket-side J padded to one line per
    # element and the partial-sum slabs, linear in the density count
    assert four - one == 3 * (two - one)
    assert two - one >= 2 * (3 + 8) * n * n * 8
    assert native_jk_workspace_bytes(object(), 1, 0.0) is None
    with pytest.raises(ValueError):
        native_jk_workspace_bytes(prov, 0, 0.0)
    with pytest.raises(ValueError):
        native_jk_workspace_bytes(prov, 1, -1.0)


def test_tuning_table_api_without_gpu():
    if not (ROOT / "gradscf_amd" / "lib" / "libgradscf_jk.so").is_file():
        pytest.skip("engine not built")
    from gradscf_amd import autotune as at
    at.reset()
    assert at.get_schedule((0, 0, 0, 0)) == at.Schedule()
    at.set_schedule(at.Schedule(chunk_items=2048, rys_lds=1), (2, 1, 1, 0))
    assert at.get_schedule((2, 1, 1, 0)).chunk_items == 2048
    assert at.get_schedule((2, 1, 1, 0)).rys_lds == 1
    assert at.get_schedule((2, 1, 1, 1)) == at.Schedule()
    assert [at.class_index(c) for c in [(0, 0, 0, 0), (1, 0, 0, 0), (2, 2, 2, 2)]] == [0, 1, 20]
    with pytest.raises(ValueError):
        at.set_schedule(at.Schedule(bras_max=0))
    with pytest.raises(ValueError):
        at.set_schedule(at.Schedule(rys_lds=2))
    at.reset()
    assert at.get_schedule((2, 1, 1, 0)) == at.Schedule()


def test_autotune_confirmation_pass_reverts_interaction_choices(monkeypatch):
    """A class whose schedule won only beside other classes' trial schedules
    (here: class A is fast at 512 threads only while B also runs 512, and B
    keeps 256) is timed again under the final table and falls back to the
    schedule it held before; a class whose choice holds keeps it."""
    if not (ROOT / "gradscf_amd" / "lib" / "libgradscf_jk.so").is_file():
        pytest.skip("engine not built")
    from gradscf_amd import autotune as at
    A, B, C = (1, 0, 0, 0), (1, 0, 1, 0), (0, 0, 0, 0)

    def fake_times(make_plan, run, reps):
        ta, tb, tc = (at.get_schedule(c).threads for c in (A, B, C))
        a = (0.8 if tb == 512 else 1.5) if ta == 512 else 1.0
        b = 1.2 if tb == 512 else 1.0
        c = 0.7 if tc == 512 else 1.0
        return {A: a, B: b, C: c}

    monkeypatch.setattr(at, "_class_times", fake_times)
    try:
        res = at.autotune(None, None, search={"threads": (256, 512)}, passes=1)
        assert res["psss"]["schedule"]["threads"] == 256 and res["psss"]["ms"] == 1.0
        assert res["psps"]["schedule"]["threads"] == 256
        assert res["ssss"]["schedule"]["threads"] == 512 and res["ssss"]["ms"] == 0.7
        assert at.get_schedule(A).threads == 256 and at.get_schedule(C).threads == 512
    finally:
        at.reset()


def test_autotune_from_a_table_never_ends_slower(monkeypatch, tmp_path):
    """Coordinate descent from the defaults settles in a local optimum here
    (threads 512, bras_max 8: 0.85) worse than a table's schedule (threads
    256, bras_max 32: 0.5); started from that table (start_from_table) the
    descent keeps it."""
    if not (ROOT / "gradscf_amd" / "lib" / "libgradscf_jk.so").is_file():
        pytest.skip("engine not built")
    from gradscf_amd import autotune as at
    A = (1, 0, 1, 0)
    landscape = {(256, 16): 1.0, (512, 16): 0.9, (512, 8): 0.85, (512, 32): 0.95, (256, 8): 0.7,
                 (256, 32): 0.5}

    def fake_times(make_plan, run, reps):
        sch = at.get_schedule(A)
        return {A: landscape[(sch.threads, sch.bras_max)]}

    monkeypatch.setattr(at, "_class_times", fake_times)
    search = {"threads": (256, 512), "bras_max": (8, 16, 32)}
    try:
        free = at.autotune(None, None, search=search, passes=1)
        assert free["psps"]["ms"] == 0.85
        at.save_table(tmp_path / "t.json", {"psps": {"schedule": at.asdict(at.Schedule(threads=256, bras_max=32))}})
        start = at.start_from_table(tmp_path / "t.json")
        assert start == {A: at.Schedule(threads=256, bras_max=32)}
        kept = at.autotune(None, None, search=search, passes=1, start=start)
        assert kept["psps"]["ms"] == 0.5 and kept["psps"]["schedule"]["bras_max"] == 32
        assert kept["psps"]["default_ms"] == 0.5
    finally:
        at.reset()


def test_engine_options_api_without_gpu():
    if not (ROOT / "gradscf_amd" / "lib" / "libgradscf_jk.so").is_file():
        pytest.skip("engine not built")
    from gradscf_amd import autotune as at
    before = at.get_options()
    default = {"general_contraction": 0, "streams": 4, "partial_sums": 1, "lds_limit": 0, "deterministic": 0,
               "balance": 1}
    assert before == default
    changed = {"general_contraction": 1, "streams": 2, "partial_sums": 1, "lds_limit": 8192, "deterministic": 1,
               "balance": 0}
    try:
        at.set_options(**changed)
        assert at.get_options() == changed
        for bad in ({"streams": 0}, {"streams": 5}, {"general_contraction": 2}, {"partial_sums": 4},
                    {"partial_sums": -1}, {"lds_limit": 100}, {"lds_limit": 200000}, {"deterministic": 2},
                    {"balance": 2}):
            with pytest.raises(ValueError):
                at.set_options(**bad)
        assert at.get_options() == changed
        with pytest.raises(ValueError):
            at.set_options(threads=4)
    finally:
        at.set_options(**before)


def test_basis_gradient_chain_rule_matches_fd():
    """The normalisation chain (env-slot gradient -> raw exponents and raw
    coefficients) against central differences of the numpy normalisation."""
    from gradscf_amd.basis import BasisParameters
    from gradscf_amd.grad import env_to_basis_gradient
    topo, par = prepare_basis(M.water_monomer(), "cc-pvdz")
    atm, bas, env = pack(topo, par)
    rng = np.random.default_rng(3)
    g = np.zeros(len(env))
    for row in bas:  # cotangent on the normalised-coefficient slots; exponent slots enter linearly
        g[row[6]:row[6] + row[2] * row[3]] = rng.normal(size=row[2] * row[3])

    def f(p):
        _, _, e = pack(topo, p)
        return float(g @ e)

    d_exp, d_coef = env_to_basis_gradient(g, bas, topo, par)
    for s in (0, 2, 4, 5):
        for p in range(len(par.exponents[s])):
            h = 1e-5 * par.exponents[s][p]
            vals = []
            for sgn in (1, -1):
                ex = list(par.exponents)
                a = np.array(ex[s], dtype=float)
                a[p] += sgn * h
                ex[s] = a
                vals.append(f(BasisParameters(tuple(ex), par.coefficients, par.centers, par.nuclear_coords)))
            assert abs((vals[0] - vals[1]) / (2 * h) - d_exp[s][p]) < 1e-7 * max(1.0, abs(d_exp[s][p]))
            c0 = np.asarray(par.coefficients[s], dtype=float)
            for col in range(c0.reshape(len(par.exponents[s]), -1).shape[1]):
                vals = []
                for sgn in (1, -1):
                    cs = list(par.coefficients)
                    c = c0.reshape(len(par.exponents[s]), -1).copy()
                    c[p, col] += sgn * 1e-6
                    cs[s] = c.reshape(c0.shape)
                    vals.append(f(BasisParameters(par.exponents, tuple(cs), par.centers, par.nuclear_coords)))
                ref = np.asarray(d_coef[s]).reshape(len(par.exponents[s]), -1)[p, col]
                assert abs((vals[0] - vals[1]) / 2e-6 - ref) < 1e-7 * max(1.0, abs(ref))


def test_padded_batch_algebra_matches_per_molecule():
    """The batched SCF's padded layout: gather/scatter round-trip the packed
    layout, and the Loewdin orthogonaliser and the per-block diagonalisation
    (molecules grouped by order, identity in the padding) give every
    molecule its own eigenvalues and occupied projector, also when only some
    molecules are solved."""
    import torch
    from gradscf_amd.batch_scf import _Padded, _diagonalize, _orthogonalizer
    from gradscf_amd.linalg import BatchedEigh
    rng = np.random.default_rng(7)
    sizes, nocc = [3, 7, 5, 7], [1, 3, 2, 2]
    offsets = np.concatenate([[0], np.cumsum(np.square(sizes))[:-1]])
    mats_f, mats_s = [], []
    for n in sizes:
        a = rng.normal(size=(n, n))
        mats_f.append(a + a.T)
        b = rng.normal(size=(n, n)) * 0.1
        mats_s.append(np.eye(n) + b @ b.T)
    packed_f = torch.as_tensor(np.concatenate([m.ravel() for m in mats_f]))
    packed_s = torch.as_tensor(np.concatenate([m.ravel() for m in mats_s]))
    B = _Padded(sizes, offsets, "cpu")
    back = torch.zeros_like(packed_f)
    B.scatter(back, B.gather(packed_f))
    assert torch.equal(back, packed_f)
    pad = 1.0 - B.valid
    S = B.gather(packed_s) + torch.diag_embed(pad)
    F = B.gather(packed_f)
    eig = BatchedEigh(sizes, "cpu")
    X = _orthogonalizer(S, eig)
    c_buf, C = _diagonalize(F, X, eig)
    fo = X.transpose(1, 2) @ F @ X
    e, _ = eig(0.5 * (fo + fo.transpose(1, 2)))
    # molecules outside the active set keep their columns
    stale = c_buf.clone()
    active = np.array([True, False, True, False])
    c2, C2 = _diagonalize(2.0 * F, X, eig, c_buf.clone(), active)
    assert torch.equal(c2[1], stale[1]) and torch.equal(c2[3], stale[3])
    assert torch.allclose(C2[0] @ C2[0].T, C[0] @ C[0].T, atol=1e-10)
    for m, (n, no) in enumerate(zip(sizes, nocc)):
        s = torch.as_tensor(mats_s[m])
        w, v = torch.linalg.eigh(s)
        x = (v * w.rsqrt()[None, :]) @ v.T
        e0, c0 = torch.linalg.eigh(x.T @ torch.as_tensor(mats_f[m]) @ x)
        c0 = x @ c0
        assert torch.allclose(e[m, :n], e0, atol=1e-10)
        co = C[m, :n, :no]
        assert torch.allclose(co @ co.T, c0[:, :no] @ c0[:, :no].T, atol=1e-10)
        if n < B.N:   # real orbitals have no padding rows
            assert float(C[m, n:, :n].abs().max()) < 1e-12


def test_out_buffer_and_packed_jk_rejections_without_gpu():
    """A caller's result buffer is checked before its pointer reaches the engine; packed
    J/K moves its operands to the GPU first, so on a host without one it fails before
    any pointer is taken."""
    import torch
    from gradscf_amd.jk import check_out_buffer, packed_jk
    dev = torch.device("cuda", 0)
    for bad in (np.zeros((2, 1, 3, 3)), torch.zeros((2, 1, 3, 3), dtype=torch.float64),
                torch.zeros((2, 1, 3, 3), dtype=torch.float32)):
        with pytest.raises(ValueError):
            check_out_buffer(bad, (2, 1, 3, 3), dev)
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError, match="needs a GPU"):
            packed_jk(torch.zeros(21, dtype=torch.float64), torch.zeros((3, 3), dtype=torch.float64))
    with pytest.raises(ValueError, match="s8"):
        packed_jk(torch.zeros((6, 6), dtype=torch.float64), torch.zeros((3, 3), dtype=torch.float64))


class _FullTensorPlan:
    """Stand-in for a device plan: J/K (and the coordinate VJP) from a full
    ERI tensor on the CPU, so the autograd rules run without a GPU."""

    def __init__(self, eri, dr_eri=None, atm=None, cutoff=0.0):
        import torch
        self.eri = torch.as_tensor(eri)
        self.dr_eri = dr_eri
        self.atm = atm
        self.cutoff = cutoff

    def jk(self, d, *, with_j=True, with_k=True):
        import torch
        out = []
        if with_j:
            out.append(torch.einsum("pqrs,...rs->...pq", self.eri, d))
        if with_k:
            out.append(torch.einsum("prqs,...rs->...pq", self.eri, d))
        return torch.stack(out)

    def jk_geometry_vjp(self, d, cj, ck):
        import torch
        # dr_eri[x] = d eri / d (coordinate x); rows = coordinates here
        g = torch.zeros(self.dr_eri.shape[0], dtype=torch.float64)
        for x in range(self.dr_eri.shape[0]):
            e = torch.as_tensor(self.dr_eri[x])
            if cj is not None:
                g[x] += torch.sum(cj * torch.einsum("pqrs,rs->pq", e, d))
            if ck is not None:
                g[x] += torch.sum(ck * torch.einsum("prqs,rs->pq", e, d))
        return g.reshape(-1, 3)


def _random_eri(n, seed):
    """A random tensor with the 8-fold permutation symmetry of real ERIs."""
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n, n, n))
    a = a + a.transpose(1, 0, 2, 3)
    a = a + a.transpose(0, 1, 3, 2)
    return a + a.transpose(2, 3, 0, 1)


def test_autodiff_density_rules_without_gpu():
    """The torch.autograd rules of direct J/K (jk.py:284-307,385-391 of the
    reference: JVP = the operator on the tangent, VJP = J on the J cotangent and
    K on the K cotangent) against autograd through explicit einsums, for the
    gradient and the Hessian-vector product, every operator selection."""
    import torch
    import torch.autograd.forward_ad as fwad
    from gradscf_amd.autodiff import JKOperator, apply_jk
    n = 5
    eri = _random_eri(n, 3)
    plan = _FullTensorPlan(eri)
    rng = np.random.default_rng(4)
    d0, t = (torch.as_tensor(x) for x in rng.normal(size=(2, 3, n, n)))
    for wj, wk in ((True, True), (True, False), (False, True)):
        op = JKOperator(plan, wj, wk)

        def loss(out):
            return 0.7 * torch.sum(out[0] ** 2) + 0.3 * torch.sum(out[-1] ** 3)

        d = d0.clone().requires_grad_(True)
        g = torch.autograd.grad(loss(apply_jk(op, d)), d, create_graph=True)[0]
        h = torch.autograd.grad(g, d, grad_outputs=t)[0]
        dr = d0.clone().requires_grad_(True)
        gr = torch.autograd.grad(loss(plan.jk(dr, with_j=wj, with_k=wk)), dr, create_graph=True)[0]
        hr = torch.autograd.grad(gr, dr, grad_outputs=t)[0]
        assert torch.allclose(g, gr, rtol=1e-12, atol=1e-9)
        assert torch.allclose(h, hr, rtol=1e-12, atol=1e-9)
        with fwad.dual_level():
            tan = fwad.unpack_dual(apply_jk(op, fwad.make_dual(d0, t))).tangent
        assert torch.allclose(tan, plan.jk(t, with_j=wj, with_k=wk), rtol=1e-12, atol=1e-12)


def test_autodiff_coordinate_rule_without_gpu():
    """Coordinate cotangents come from the plan's geometry VJP summed over the
    density batch; a screened operator refuses, and a second derivative
    through a coordinate gradient raises instead of returning zeros."""
    import torch
    from gradscf_amd.autodiff import JKOperator, apply_jk
    n = 4
    eri = _random_eri(n, 5)
    dr = np.stack([_random_eri(n, 10 + x) for x in range(6)])      # two rows of three coordinates
    plan = _FullTensorPlan(eri, dr_eri=dr, atm=np.zeros((2, 6), dtype=np.int32))
    rng = np.random.default_rng(6)
    d = torch.as_tensor(rng.normal(size=(2, n, n)))
    cj, ck = (torch.as_tensor(x) for x in rng.normal(size=(2, 2, n, n)))
    coords = torch.zeros((2, 3), dtype=torch.float64, requires_grad=True)
    out = apply_jk(JKOperator(plan, True, True), d, coords)
    (torch.sum(cj * out[0]) + torch.sum(ck * out[1])).backward()
    ref = sum(plan.jk_geometry_vjp(d[b], cj[b], ck[b]) for b in range(2))
    assert torch.allclose(coords.grad, ref, rtol=1e-12, atol=1e-12)
    coords = torch.zeros((2, 3), dtype=torch.float64, requires_grad=True)
    out = apply_jk(JKOperator(plan, True, True), d, coords)
    g = torch.autograd.grad(torch.sum(out[0] ** 2), coords, create_graph=True)[0]
    with pytest.raises(NotImplementedError, match="Second derivatives"):
        torch.autograd.grad(g.sum(), coords)
    screened = _FullTensorPlan(eri, dr_eri=dr, atm=plan.atm, cutoff=1e-12)
    coords = torch.zeros((2, 3), dtype=torch.float64, requires_grad=True)
    out = apply_jk(JKOperator(screened, True, False), d, coords)
    with pytest.raises(NotImplementedError, match="cutoff=0"):
        out.sum().backward()


def test_differentiable_packing_matches_pack():
    """packing.pack_torch gives pack's environment, and its gradient chains a
    coefficient-slot cotangent to the raw coefficients like grad.env_to_basis_gradient."""
    import torch
    from gradscf_amd.basis import BasisParameters
    from gradscf_amd.grad import env_to_basis_gradient
    from gradscf_amd.packing import pack_torch
    topo, par = prepare_basis(M.water_monomer(), "cc-pvdz", cart=False)
    atm, bas, env = pack(topo, par)
    exps = tuple(torch.tensor(np.asarray(a, dtype=float), requires_grad=True) for a in par.exponents)
    coefs = tuple(torch.tensor(np.asarray(c, dtype=float), requires_grad=True) for c in par.coefficients)
    atm2, bas2, env_t = pack_torch(topo, BasisParameters(exps, coefs, par.centers, par.nuclear_coords))
    assert np.array_equal(atm, atm2) and np.array_equal(bas, bas2)
    assert np.allclose(env_t.detach().numpy(), env, rtol=1e-14, atol=0.0)
    w = np.random.default_rng(7).normal(size=env.shape)
    (torch.as_tensor(w) * env_t).sum().backward()
    d_exp, d_coef = env_to_basis_gradient(w, bas, topo, par)
    for s in range(len(exps)):
        assert np.allclose(exps[s].grad.numpy(), d_exp[s], rtol=1e-12, atol=1e-12)
        assert np.allclose(coefs[s].grad.numpy(), d_coef[s], rtol=1e-12, atol=1e-12)
