"""The allocation-free J/K build (gscf_jk_build_noalloc, workspace in the
library's file-scope device scratch array) against gscf_jk_build through the
C ABI: two calls in a row, the second with a smaller input, results equal to
FP64 rounding (1e-12) and to the oracle (1e-10); a request larger than the
scratch array is rejected before any launch."""
import ctypes

import numpy as np
import pytest

from tests import systems

pytestmark = pytest.mark.gpu


def _call(fn, plan, d, nops, with_j, with_k, flags):
    import torch
    from gradscf_amd import _native
    nb = d.shape[0]
    out = torch.full((nops, nb, plan.nao, plan.nao), float("nan"), dtype=torch.float64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(plan._handle, d.data_ptr(), nb, out.data_ptr(), with_j, with_k, flags, ctypes.c_double(0.0), stream)
    return rc, out


def test_noalloc_build_matches_build(oracle_mod):
    import torch
    from gradscf_amd import _native
    from gradscf_amd.jk import DirectJKPlan
    L = _native.lib()
    _, _, atm, bas, env = systems.tables(systems.water_dimer(), "cc-pvdz", cart=False)
    plan = DirectJKPlan(atm, bas, env, cart=False, cutoff=1e-12)
    big = systems.random_density(plan.nao, nb=2, seed=901, symmetric=False)
    small = systems.random_density(plan.nao, nb=1, seed=902, symmetric=True)
    cases = [(big, 0, 1, 1), (small, _native.FLAG_SYMMETRIC_DENSITY, 1, 1), (small, 0, 0, 1)]
    for d, flags, wj, wk in cases:
        dd = torch.as_tensor(d, device="cuda").contiguous()
        nops = wj + wk
        rc, got = _call(L.gscf_jk_build_noalloc, plan, dd, nops, wj, wk, flags)
        _native.check(rc)
        rc, ref = _call(L.gscf_jk_build, plan, dd, nops, wj, wk, flags)
        _native.check(rc)
        torch.cuda.synchronize()
        assert float((got - ref).abs().max()) < 1e-12
        oref = oracle_mod.direct_jk(atm, bas, env, d, cart=False, cutoff=1e-12, workers=4,
                                    with_j=bool(wj), with_k=bool(wk))
        assert np.max(np.abs(got.cpu().numpy() - oref)) < 1e-10


def test_noalloc_build_rejects_requests_beyond_its_scratch():
    import torch
    from gradscf_amd import _native
    from gradscf_amd import molecule as M
    from gradscf_amd.jk import DirectJKPlan
    L = _native.lib()
    _, _, atm, bas, env = systems.tables(M.water_cluster(16, seed=16, box=12.0), "cc-pvdz", cart=True)
    plan = DirectJKPlan(atm, bas, env, cart=True, cutoff=1e-12)
    # 400 Cartesian AO: 4 nonsymmetric densities need 8 * 160000 * (24 + 32) bytes > 64 MiB
    d = torch.zeros((4, plan.nao, plan.nao), dtype=torch.float64, device="cuda")
    rc, _ = _call(L.gscf_jk_build_noalloc, plan, d, 2, 1, 1, 0)
    assert rc == _native.GSCF_INVALID_ARGUMENT
    assert b"scratch" in L.gscf_last_error()
    # one nonsymmetric density fits and matches the allocating build
    d1 = torch.as_tensor(systems.random_density(plan.nao, nb=1, seed=903, symmetric=False) * 0.05,
                         device="cuda").contiguous()
    rc, got = _call(L.gscf_jk_build_noalloc, plan, d1, 2, 1, 1, 0)
    _native.check(rc)
    rc, ref = _call(L.gscf_jk_build, plan, d1, 2, 1, 1, 0)
    _native.check(rc)
    assert float((got - ref).abs().max()) < 1e-12
