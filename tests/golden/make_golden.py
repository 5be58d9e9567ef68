"""Regenerate the golden J/K fixtures (run here, where oracle/_ref exists).

Every fixture is computed by the oracle's restatement of the reference DirectJK
driver running on the reference's own libcint (oracle/_ref/libcint_ref.so,
compiled from the reference sources), i.e. the reference arithmetic:

  water_631g_jk.npz        water / 6-31G (13 AO, Cartesian), 2 nonsymmetric densities, cutoff 0
  dimer_ccpvdz_sph_jk.npz  (H2O)2 seed 7 / cc-pVDZ spherical (48 AO), symmetric density, cutoff 1e-12
  water16_ccpvdz_sph_sample.npz
                           (H2O)16 seed 16 / cc-pVDZ spherical (384 AO), D = C C^T with
                           C = default_rng(16).normal(size=(384, 80)) / sqrt(384), cutoff
                           1e-12: 4000 sampled J and K elements,
                           tr(DJ), tr(DK) and Frobenius norms
  scf_energies.json        oracle RHF energies: water / 6-31G (Cartesian, the config-1
                           geometry, cutoff 0), and with --large the reference-libcint
                           FP64 RHF energies of C20H42 (all-trans) and (H2O)32 (seed 32,
                           15 A cube) in 6-31G* Cartesian and (H2O)16 (seed 16, 12 A cube)
                           in cc-pVDZ spherical, Schwarz cutoff 1e-12

Usage: python tests/golden/make_golden.py [--large]
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

import oracle  # noqa: E402
from oracle import scf as oscf  # noqa: E402
from gradscf_amd import molecule as M  # noqa: E402
from tests import systems  # noqa: E402


def main():
    assert oracle.have_ref(), "build oracle/_ref first (make -C oracle ref)"
    rng = np.random.default_rng(2024)
    with oracle.backend("libcint"):
        _, _, atm, bas, env = systems.tables(M.water_monomer(), "6-31g", cart=True)
        d = rng.normal(size=(2, 13, 13))
        out = oracle.direct_jk(atm, bas, env, d, cart=True, cutoff=0.0, workers=2)
        np.savez_compressed(HERE / "water_631g_jk.npz", density=d, J=out[0], K=out[1])

        _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=False)
        n = oracle.nao(bas, False)
        d = rng.normal(size=(n, n))
        d = 0.5 * (d + d.T)
        out = oracle.direct_jk(atm, bas, env, d, cart=False, cutoff=1e-12, workers=8)
        np.savez_compressed(HERE / "dimer_ccpvdz_sph_jk.npz", density=d, J=out[0, 0], K=out[1, 0])

        spec = M.water_cluster(16, seed=16, box=12.0)
        _, _, atm, bas, env = systems.tables(spec, "cc-pvdz", cart=False)
        n = oracle.nao(bas, False)
        c = np.random.default_rng(16).normal(size=(n, 80)) / np.sqrt(n)
        d = c @ c.T
        out = oracle.direct_jk(atm, bas, env, d, cart=False, cutoff=1e-12, workers=8)
        idx = np.random.default_rng(17).integers(0, n, size=(4000, 2))
        J, K = out[0, 0], out[1, 0]
        np.savez_compressed(HERE / "water16_ccpvdz_sph_sample.npz", idx=idx,
                            J=J[idx[:, 0], idx[:, 1]], K=K[idx[:, 0], idx[:, 1]],
                            trDJ=np.sum(d * J), trDK=np.sum(d * K),
                            normJ=np.linalg.norm(J), normK=np.linalg.norm(K))

    spec = M.water_monomer()
    _, _, atm, bas, env = systems.tables(spec, "6-31g", cart=True)
    s, t, v = oracle.int1e(atm, bas, env, True)

    def jk(dm):
        o = oracle.direct_jk(atm, bas, env, dm, cart=True)
        return o[0, 0], o[1, 0]

    r = oscf.rhf(t + v, s, spec.nelectron, spec.nuclear_repulsion(), jk)
    (HERE / "scf_energies.json").write_text(json.dumps(
        {"water_631g_cart": {"energy": r["energy"], "cycles": r["cycles"], "converged": r["converged"],
                             "geometry": "O 0 0 0; H 0 0.757 0.587; H 0 -0.757 0.587 (Angstrom)"}},
        indent=1))


LARGE = {
    # key: (molecule, basis, cart)
    "c20h42_631gs_cart": (lambda: M.linear_alkane(20), "6-31g*", True),
    "water32_631gs_cart": (lambda: M.water_cluster(32, seed=32, box=15.0), "6-31g*", True),
    "water16_ccpvdz_sph": (lambda: M.water_cluster(16, seed=16, box=12.0), "cc-pvdz", False),
}


def large(workers: int = 8):
    """FP64 RHF energies of the larger workloads on the reference libcint (cutoff 1e-12)."""
    import time
    path = HERE / "scf_energies.json"
    table = json.loads(path.read_text())
    for key, (make, basis, cart) in LARGE.items():
        spec = make()
        _, _, atm, bas, env = systems.tables(spec, basis, cart=cart)
        with oracle.backend("libcint"):
            s, t, v = oracle.libcint_int1e(atm, bas, env, cart)

            def jk(dm):
                o = oracle.direct_jk(atm, bas, env, dm, cart=cart, cutoff=1e-12, workers=workers)
                return o[0, 0], o[1, 0]

            t0 = time.perf_counter()
            r = oscf.rhf(t + v, s, spec.nelectron, spec.nuclear_repulsion(), jk)
        print(f"{key}: E={r['energy']:.12f} cycles={r['cycles']} converged={r['converged']} "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
        table[key] = {"energy": r["energy"], "cycles": r["cycles"], "converged": r["converged"],
                      "basis": basis, "cart": cart, "cutoff": 1e-12, "nao": int(s.shape[0]),
                      "eri": "reference libcint 6.1.3 (oracle/_ref)"}
        path.write_text(json.dumps(table, indent=1))


if __name__ == "__main__":
    if "--large" in sys.argv[1:]:
        large()
    else:
        main()
