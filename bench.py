#!/usr/bin/env python
"""Benchmark of the MI355X direct-SCF J/K engine.

Workload (one "step" = one full direct J/K build): (H2O)16, cc-pVDZ, spherical
(384 AO, 176 shells), Schwarz cutoff 1e-12, density = the converged FP64 RHF
density of the same molecule (computed here by the GPU SCF, whose wall time is
reported too).  Geometry: rigid waters (O-H 0.96 A, HOH 104.5 deg), random
orientations, O positions uniform in a 12 A cube with O-O >= 2.6 A, seed 16.

Prints ONE JSON line: J/K builds/s (value), ms per build, SCF wall time,
per-class kernel roofline of the dominant class kernel (algorithmic FLOPs from
gradscf_amd.roofline over the HIP-event durations of the timed builds), and a
CPU baseline: the reference's libcint (compiled from the reference sources
into oracle/_ref) driven by the oracle's restatement of the reference
DirectJK, timed on this host's cores on the same density.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix) published peak
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(atm, bas, env, dens, cutoff, builds, j_gpu, k_gpu, scf_job=None):
    """Reference libcint J/K (restated DirectJK driver, row-ownership workers) on
    the host cores: `builds` builds with every core after a warm-up, one build
    with one worker, and (scf_job = (spec, gpu_energy)) the full RHF of the
    same workload with every core, from integrals ready to convergence."""
    import numpy as np
    import oracle
    from oracle import scf as oscf
    cores = len(os.sched_getaffinity(0))
    kind = "reference" if oracle.have_ref() else "port"
    backend = "libcint" if kind == "reference" else "oracle"
    times = []
    scf_rec = None
    with oracle.backend(backend):
        for i in range(builds + 1):
            t0 = time.perf_counter()
            out = oracle.direct_jk(atm, bas, env, dens, cart=False, cutoff=cutoff, workers=cores)
            dt = time.perf_counter() - t0
            log(f"cpu baseline build {i}: {dt:.3f} s ({backend}, {cores} workers)")
            if i:
                times.append(dt)
        t0 = time.perf_counter()
        oracle.direct_jk(atm, bas, env, dens, cart=False, cutoff=cutoff, workers=1)
        one = time.perf_counter() - t0
        log(f"cpu baseline build, 1 worker: {one:.3f} s")
        if scf_job is not None:
            spec, e_gpu = scf_job
            s, t, v = oracle.libcint_int1e(atm, bas, env, False) if kind == "reference" else \
                oracle.int1e(atm, bas, env, False)

            def jk(dm):
                o = oracle.direct_jk(atm, bas, env, dm, cart=False, cutoff=cutoff, workers=cores)
                return o[0, 0], o[1, 0]

            t0 = time.perf_counter()
            r = oscf.rhf(t + v, s, spec.nelectron, spec.nuclear_repulsion(), jk)
            scf_s = time.perf_counter() - t0
            log(f"cpu baseline SCF: {scf_s:.1f} s, {r['cycles']} cycles, E {r['energy']:.10f} "
                f"(GPU - CPU {e_gpu - r['energy']:.2e})")
            scf_rec = {"wall_s": scf_s, "cycles": r["cycles"], "converged": bool(r["converged"]),
                       "energy": r["energy"], "energy_diff_gpu_minus_cpu": e_gpu - r["energy"], "workers": cores}
    med = statistics.median(times)
    err_j = float(np.max(np.abs(out[0, 0] - j_gpu)))
    err_k = float(np.max(np.abs(out[1, 0] - k_gpu)))
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": 1.0 / med, "unit": "J/K builds/s", "cores": cores, "kind": kind,
            "sample": f"{builds} full J/K builds of the same workload (median, after 1 warm-up), "
                      f"Schwarz bounds + pair data rebuilt per call as the reference does; {cpu_model}",
            "ms_per_build": med * 1e3, "ms_per_build_1_worker": one * 1e3,
            "max_abs_diff_vs_gpu": {"J": err_j, "K": err_k}, "scf": scf_rec}


CLASS_LETTERS = "spdf"
HBM_PROFILE = os.path.join(ROOT, "profiles", "jk_class_kernels_water16_ccpvdz_hbm.txt")
TUNING_DIR = os.path.join(ROOT, "gradscf_amd", "tuning")


def _kernel_class(name: str):
    """'void gscf::jk_bra_kernel<2, 1, 1, 0, ...>(gscf::BraArgs)' -> ('jk_bra_kernel', 'dpps');
    the in-core variants of the bra kernel (MODE template argument 1: evaluate
    and store, 2: digest stored blocks) get '/store' or '/load', the
    deterministic-mode variant '/det'."""
    import re
    m = re.search(r"(\w+)<(\d), (\d), (\d), (\d)((?:, \w+)*)>", name)
    if not m:
        return None, None
    cls = "".join(CLASS_LETTERS[int(v)] for v in m.groups()[1:5])
    rest = [x.strip() for x in m.group(6).split(",") if x.strip()]
    if m.group(1) == "jk_bra_kernel" and len(rest) >= 7:
        # rest: column counts (4), Rys table in LDS, mixed precision, MODE[, DET]
        if rest[6] in ("1", "2"):
            cls += "/store" if rest[6] == "1" else "/load"
        if len(rest) >= 8 and rest[7] in ("1", "true"):
            cls += "/det"
    return m.group(1), cls


def copy_calibration() -> int:
    """A device copy of known size (256 MiB read + 256 MiB written, 3 times):
    the reference reading for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 1 << 25
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)  # one elementwise kernel: n*8 bytes read, n*8 written
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n * 8, "launches": 3}), flush=True)
    return 0


def measured_peaks() -> dict:
    """Achievable FP64 FMA rate (the library's probe: 8 independent FMA chains
    per lane) and device copy bandwidth (1 GiB read + 1 GiB written per copy)
    on this GPU, quoted beside the published peaks the roofline uses."""
    import ctypes
    import torch
    from gradscf_amd import _native
    blocks, iters = 256 * 16, 1 << 14
    out = torch.empty(blocks * 256, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()

    def probe():
        _native.check(_native.lib().gscf_fma_probe(blocks, iters, ctypes.c_void_p(out.data_ptr()),
                                                   ctypes.c_void_p(stream.cuda_stream)))

    def timed(f, reps):
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    t = timed(probe, 3)
    fma_tflops = blocks * 256 * 16 * iters / t / 1e12
    n = 1 << 27
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    t = timed(lambda: b.copy_(a), 5)
    copy_gbs = 2 * n * 8 / t / 1e9
    del a, b, out
    return {"fp64_fma_tflops": fma_tflops, "copy_gbs": copy_gbs,
            "published": {"fp64_tflops": FP64_PEAK_TFLOPS, "hbm_gbs": 8000.0},
            "method": "FMA: 4096 x 256 lanes x 8 chains x 16384 steps (library probe); copy: torch copy_ of "
                      "1 GiB FP64, read + write bytes counted; HIP events, after a warm-up"}


def summarize_profile(paths) -> int:
    """Summarise rocprofv3 CSV output (--output-format csv): per-kernel
    statistics from *kernel_stats.csv and, from *counter_collection.csv, the
    average of every counter per dispatch of each kernel."""
    import csv
    import glob
    for path in paths:
        files = sorted(glob.glob(os.path.join(path, "**", "*.csv"), recursive=True))
        print(f"# {path}")
        for f in files:
            if f.endswith("kernel_stats.csv"):
                rows = list(csv.DictReader(open(f)))
                print(f"{'kernel':72s} {'calls':>6s} {'total_ms':>9s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'pct':>6s}")
                for r in rows:
                    kind, cls = _kernel_class(r["Name"])
                    label = f"{kind}[{cls}]" if kind else r["Name"][:72]
                    print(f"{label:72s} {int(r['Calls']):6d} {float(r['TotalDurationNs']) / 1e6:9.3f} "
                          f"{float(r['AverageNs']) / 1e3:9.2f} {float(r['MinNs']) / 1e3:9.2f} "
                          f"{float(r['MaxNs']) / 1e3:9.2f} {float(r['Percentage']):6.2f}")
            elif f.endswith("counter_collection.csv"):
                acc = {}
                for r in csv.DictReader(open(f)):
                    key = (r["Kernel_Name"], r["Counter_Name"])
                    s = acc.setdefault(key, [0.0, set()])
                    s[0] += float(r["Counter_Value"])
                    s[1].add(r["Dispatch_Id"])
                by = {}
                for (k, c), (v, ds) in acc.items():
                    by.setdefault(k, {})[c] = (v / len(ds), len(ds))
                print("# counters: average per dispatch (n = dispatches)")
                for k in sorted(by):
                    kind, cls = _kernel_class(k)
                    label = f"{kind}[{cls}]" if kind else k[:60]
                    n = max(d for _, d in by[k].values())
                    vals = " ".join(f"{c}={v:.6g}" for c, (v, _) in sorted(by[k].items()))
                    print(f"{label:40s} n={n:<5d} {vals}")
    return 0


def traffic_table(fetch_dir, write_dir, cal_fetch_dir, cal_write_dir) -> int:
    """HBM bytes per launch of every J/K class kernel: FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) from two counter runs of the workload, each
    scaled by bytes / counter of the copy calibration kernel; printed as a
    table and as one TRAFFIC_JSON line."""
    import csv
    import glob

    def per_kernel(path, counter):
        acc = {}
        for f in glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] != counter:
                    continue
                s = acc.setdefault(r["Kernel_Name"], {})
                s[r["Dispatch_Id"]] = s.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
        return {k: sum(v.values()) / len(v) for k, v in acc.items()}

    copy_bytes = (1 << 25) * 8
    factors = {}
    for counter, path in (("FETCH_SIZE", cal_fetch_dir), ("WRITE_SIZE", cal_write_dir)):
        vals = [v for k, v in per_kernel(path, counter).items() if "elementwise" in k]
        if not vals:
            print(f"no calibration kernel in {path}")
            return 1
        factors[counter] = copy_bytes / (max(vals) * 1024.0)
    fetch = per_kernel(fetch_dir, "FETCH_SIZE")
    write = per_kernel(write_dir, "WRITE_SIZE")
    table = {}
    print(f"calibration (bytes per counter byte): FETCH_SIZE {factors['FETCH_SIZE']:.4f}, "
          f"WRITE_SIZE {factors['WRITE_SIZE']:.4f}")
    print(f"{'kernel':32s} {'fetch_MB':>10s} {'write_MB':>10s} {'total_MB':>10s}")
    for k in sorted(fetch):
        kind, cls = _kernel_class(k)
        if kind not in ("jk_bra_kernel", "jk_full_kernel", "jk_split_kernel"):
            continue
        fb = fetch[k] * 1024.0 * factors["FETCH_SIZE"]
        wb = write.get(k, 0.0) * 1024.0 * factors["WRITE_SIZE"]
        table[cls] = {"kernel": kind, "fetch_bytes": fb, "write_bytes": wb, "bytes_per_launch": fb + wb}
        print(f"{kind + '[' + cls + ']':32s} {fb / 1e6:10.3f} {wb / 1e6:10.3f} {(fb + wb) / 1e6:10.3f}")
    print("TRAFFIC_JSON " + json.dumps({"factors": factors, "classes": table}))
    return 0


def profiled_traffic(cls: str):
    """HBM bytes per launch of the class kernel from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE, KiB per dispatch, scaled by the copy
    calibration factor recorded in the same file); None when absent."""
    try:
        text = open(HBM_PROFILE).read()
    except OSError:
        return None
    for line in text.splitlines():
        if line.startswith("TRAFFIC_JSON "):
            table = json.loads(line[len("TRAFFIC_JSON "):])
            rec = table.get("classes", {}).get(cls)
            return None if rec is None else rec["bytes_per_launch"]
    return None


def batch_section(thr: float, reps: int = 5, autotune: bool = True, retune: bool = False, save_dir=None,
                  incore: bool = True) -> dict:
    """256 seeded water clusters (H2O)_n, n ~ U{1..8} (default_rng(256), cluster
    i seeded 1000+i, 8 A cube), 6-31G* Cartesian, Schwarz cutoff 1e-12: batched
    FP64 and mixed SCF with every iteration's J/K for all molecules in one
    graph replay; J/K throughput at the converged densities, batched mixed vs
    batched FP64 vs one plain FP64 call per molecule."""
    import numpy as np
    import torch
    from gradscf_amd import molecule as M
    from gradscf_amd.basis import prepare_basis
    from gradscf_amd.batch_scf import _GraphJK, rhf_batch
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.packing import pack

    specs = M.batch_water_clusters(256, seed=256, base_seed=1000, nmax=8, box=8.0)
    tabs = [pack(*prepare_basis(s, "6-31g*", cart=True)) for s in specs]
    nel = [s.nelectron for s in specs]
    enuc = [s.nuclear_repulsion() for s in specs]
    from gradscf_amd import autotune as at
    tuning = None
    if autotune:
        table_path = os.path.join(TUNING_DIR, "batch256_631gs_cart.json")
        if os.path.isfile(table_path) and not retune:
            tuning = {"source": os.path.relpath(table_path, ROOT), "table": at.load_table(table_path)}
        else:
            at.reset()
            p0 = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
            d0 = rhf_batch(p0, nel, enuc, use_graph=False, max_cycle=3).density
            del p0
            result = at.autotune(lambda: DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12),
                                 lambda p: p.jk(d0[None], symmetric=True, validate=False), reps=2,
                                 search={"rys_lds": (0, 1), "ket_lds": (0, 1), "threads": (256, 512),
                                         ("chunk_items", "target_prim"): tuple(
                                             (c, t) for c in (2048, 4096, 8192) for t in (24, 48, 96)),
                                         "bras_max": (2, 4, 8, 16, 32),
                                         "lds_pack": (8192, 16384, 32768, 49152, 65536)},
                                 log=log)
            tuning = {"source": "autotuned in this run", "classes": result}
            if save_dir:
                at.save_table(os.path.join(save_dir, "batch256_631gs_cart.json"), result)
    else:
        at.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = DirectJKPlan.batch(tabs, cart=True, cutoff=1e-12)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    log(f"batch: {plan.nmol} molecules, {int(plan.mol_n.sum())} AO, {plan.info['nquartets']} quartets, "
        f"plan {plan_s * 1e3:.0f} ms")
    # the first SCF of the process also pays one-time start-up (graph capture,
    # first solver calls at these shapes): reported as wall_s_first
    ref_first = rhf_batch(plan, nel, enuc, use_graph=True)
    ref = rhf_batch(plan, nel, enuc, use_graph=True)
    log(f"batch FP64 SCF: {ref.timings['scf_s']:.3f} s (first in process {ref_first.timings['scf_s']:.3f} s), "
        f"{len(ref.iteration_s)} iterations, converged {int(ref.converged.sum())}/{plan.nmol}")
    mix = rhf_batch(plan, nel, enuc, use_graph=True, mixed_threshold=thr)
    err = np.abs(mix.energies - ref.energies)
    it_ms = np.asarray(mix.iteration_s) * 1e3
    log(f"batch mixed SCF: {mix.timings['scf_s']:.3f} s, {len(mix.iteration_s)} iterations, "
        f"max |dE| {err.max():.2e}, p50 {np.percentile(it_ms, 50):.2f} ms p99 {np.percentile(it_ms, 99):.2f} ms")

    # J/K throughput at the converged densities
    dens = ref.density

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps * 1e3

    g64 = _GraphJK(plan, None)
    gmx = _GraphJK(plan, thr)
    ms64 = timed(lambda: g64(dens))
    msmx = timed(lambda: gmx(dens))
    j64 = g64(dens)
    jmx = gmx(dens)
    jk_err = max(float((jmx[0] - j64[0]).abs().max()), float((jmx[1] - j64[1]).abs().max()))
    singles = [DirectJKPlan(a, b, e, cart=True, cutoff=1e-12) for a, b, e in tabs]
    parts = [d[None].contiguous() for d in plan.unpack(dens)]
    outs = [torch.empty((2, 1, p.nao, p.nao), dtype=torch.float64, device="cuda") for p in singles]

    def per_molecule():
        for p, d, o in zip(singles, parts, outs):
            p.jk(d, symmetric=True, validate=False, out=o)

    ms_single = timed(per_molecule)
    log(f"batch J/K: FP64 batched {ms64:.2f} ms, mixed batched {msmx:.2f} ms, FP64 per-molecule calls "
        f"{ms_single:.2f} ms (mixed batched speedup {ms_single / msmx:.2f}x vs plain per-molecule, "
        f"{ms64 / msmx:.2f}x vs FP64 batched)")
    inc_rec = None
    if incore:
        nbytes = plan.incore_bytes()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.incore(True)
        torch.cuda.synchronize()
        store_ms = (time.perf_counter() - t0) * 1e3
        gin = _GraphJK(plan, None)
        msin = timed(lambda: gin(dens))
        jin = gin(dens)
        in_err = max(float((jin[0] - j64[0]).abs().max()), float((jin[1] - j64[1]).abs().max()))
        inc = rhf_batch(plan, nel, enuc, use_graph=True)
        plan.incore(False)
        ierr = np.abs(inc.energies - ref.energies)
        iit = np.asarray(inc.iteration_s) * 1e3
        log(f"batch in-core: {nbytes / 1e9:.2f} GB stored in {store_ms:.1f} ms, J/K {msin:.2f} ms "
            f"({ms_single / msin:.2f}x plain per-molecule), SCF {inc.timings['scf_s']:.3f} s, "
            f"p50 {np.percentile(iit, 50):.2f} ms, max |dE| {ierr.max():.2e}")
        inc_rec = {"bytes": nbytes, "store_ms": store_ms, "jk_ms": msin,
                   "speedup_vs_fp64_per_molecule": ms_single / msin, "speedup_vs_fp64_batched": ms64 / msin,
                   "max_abs_jk_diff_vs_fp64": in_err,
                   "scf": {"wall_s": inc.timings["scf_s"], "wall_s_including_store": inc.timings["scf_s"]
                           + store_ms * 1e-3, "molecules_per_s": plan.nmol / inc.timings["scf_s"],
                           "iterations": len(inc.iteration_s), "converged": int(inc.converged.sum()),
                           "iteration_ms_p50": float(np.percentile(iit, 50)),
                           "iteration_ms_p99": float(np.percentile(iit, 99)),
                           "energy_err_vs_fp64_max": float(ierr.max())}}
    return {
        "incore": inc_rec,
        "workload": "256 water clusters (H2O)_n, n~U{1..8} (default_rng(256), cluster i seed 1000+i, 8 A cube), "
                    "6-31G* Cartesian, Schwarz 1e-12",
        "molecules": int(plan.nmol), "ao_total": int(plan.mol_n.sum()), "quartets": int(plan.info["nquartets"]),
        "tuning": tuning,
        "plan_ms": plan_s * 1e3, "mixed_threshold": thr,
        "scf_mixed": {"wall_s": mix.timings["scf_s"], "molecules_per_s": plan.nmol / mix.timings["scf_s"],
                      "iterations": len(mix.iteration_s), "converged": int(mix.converged.sum()),
                      "iteration_ms_p50": float(np.percentile(it_ms, 50)),
                      "iteration_ms_p99": float(np.percentile(it_ms, 99)),
                      "energy_err_vs_fp64_max": float(err.max()), "energy_err_vs_fp64_mean": float(err.mean())},
        "scf_fp64": {"wall_s": ref.timings["scf_s"], "wall_s_first": ref_first.timings["scf_s"],
                     "molecules_per_s": plan.nmol / ref.timings["scf_s"],
                     "iterations": len(ref.iteration_s), "converged": int(ref.converged.sum()),
                     "iteration_ms_p50": float(np.percentile(np.asarray(ref.iteration_s) * 1e3, 50))},
        "jk": {"fp64_batched_ms": ms64, "mixed_batched_ms": msmx, "fp64_per_molecule_calls_ms": ms_single,
               "mixed_batched_speedup_vs_fp64_per_molecule": ms_single / msmx,
               "mixed_batched_speedup_vs_fp64_batched": ms64 / msmx,
               "mixed_vs_fp64_max_abs_jk_diff": jk_err},
    }


def water_section() -> dict:
    """Water (O 0 0 0; H 0 +-0.757 0.587 A), 6-31G Cartesian, core guess, DIIS,
    conv_tol 1e-10: one J/K build and the full SCF on the GPU (the CPU
    reference for the same workload is timed in the cpu_baseline leg)."""
    import torch
    from gradscf_amd import molecule as M
    from gradscf_amd.basis import prepare_basis
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.packing import pack
    from gradscf_amd.scf import rhf

    spec = M.water_monomer()
    atm, bas, env = pack(*prepare_basis(spec, "6-31g", cart=True))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = DirectJKPlan(atm, bas, env, cart=True, cutoff=0.0)
    torch.cuda.synchronize()
    plan_ms = (time.perf_counter() - t0) * 1e3
    res = rhf(plan, spec.nelectron, spec.nuclear_repulsion(), conv_tol=1e-10)
    d = res.density.contiguous()
    out = torch.empty((2, 1, plan.nao, plan.nao), dtype=torch.float64, device="cuda")
    for _ in range(5):
        plan.jk(d[None], symmetric=True, validate=False, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(100):
        plan.jk(d[None], symmetric=True, validate=False, out=out)
    torch.cuda.synchronize()
    gpu_ms = (time.perf_counter() - t0) / 100 * 1e3
    log(f"water 6-31G: GPU J/K {gpu_ms * 1e3:.1f} us/build, SCF {res.timings['scf_s']:.3f} s, E {res.energy:.10f}")
    return {"workload": "water (O 0 0 0; H 0 +-0.757 0.587 A), 6-31G Cartesian, 13 AO",
            "tables": (atm, bas, env), "spec": spec, "density": d.cpu().numpy(), "j": out[0, 0].cpu().numpy(),
            "k": out[1, 0].cpu().numpy(),
            "record": {"energy": res.energy, "gpu_jk_ms": gpu_ms, "gpu_plan_ms": plan_ms,
                       "gpu_scf_s": res.timings["scf_s"], "gpu_scf_cycles": res.cycles}}


def cpu_baseline_water(w: dict) -> dict:
    """The reference CPU path on the water/6-31G workload: libcint J/K through
    the restated DirectJK (one worker) and the restated RHF loop."""
    import numpy as np
    import oracle
    from oracle import scf as oscf
    atm, bas, env = w["tables"]
    spec, dh = w["spec"], w["density"]
    backend = "libcint" if oracle.have_ref() else "oracle"
    with oracle.backend(backend):
        oracle.direct_jk(atm, bas, env, dh, cart=True, cutoff=0.0, workers=1)
        t0 = time.perf_counter()
        for _ in range(20):
            ref = oracle.direct_jk(atm, bas, env, dh, cart=True, cutoff=0.0, workers=1)
        cpu_ms = (time.perf_counter() - t0) / 20 * 1e3
        s, t, v = oracle.int1e(atm, bas, env, True)

        def jk(dm):
            o = oracle.direct_jk(atm, bas, env, dm, cart=True, cutoff=0.0, workers=1)
            return o[0, 0], o[1, 0]

        t0 = time.perf_counter()
        ores = oscf.rhf(t + v, s, spec.nelectron, spec.nuclear_repulsion(), jk)
        cpu_scf_s = time.perf_counter() - t0
    err = float(max(np.max(np.abs(ref[0, 0] - w["j"])), np.max(np.abs(ref[1, 0] - w["k"]))))
    log(f"water 6-31G CPU reference: J/K {cpu_ms:.3f} ms/build, SCF {cpu_scf_s:.3f} s, "
        f"dE(GPU-CPU) {w['record']['energy'] - ores['energy']:.2e}")
    return {"kind": "reference" if backend == "libcint" else "port", "cores": 1, "jk_ms": cpu_ms,
            "scf_s": cpu_scf_s, "scf_cycles": ores.get("cycles"), "energy": ores["energy"],
            "energy_diff_gpu_minus_cpu": w["record"]["energy"] - ores["energy"],
            "jk_max_abs_diff_vs_gpu": err}


def mixed_section(thr: float, tuned: bool = True) -> dict:
    """C20H42 (all-trans) and (H2O)32 (seed 32, 15 A cube), 6-31G* Cartesian,
    Schwarz 1e-12: full RHF in FP64 and in mixed precision (FP32 for quartets
    with Q_ij Q_kl max|D| < thr), energy error and J/K build times."""
    from gradscf_amd import autotune as at
    # schedules tuned on (H2O)32 6-31G* Cartesian (used for both molecules);
    # the headline table is installed again afterwards
    c3_table = os.path.join(TUNING_DIR, "water32_631gs_cart.json")
    tuned = tuned and os.path.isfile(c3_table)
    out = {"tuning": os.path.relpath(c3_table, ROOT) if tuned else None}
    if tuned:
        at.load_table(c3_table)
    try:
        _mixed_molecules(thr, out)
    finally:
        head_table = os.path.join(TUNING_DIR, "water16_ccpvdz_sph.json")
        if tuned and os.path.isfile(head_table):
            at.load_table(head_table)
    return out


def _mixed_molecules(thr: float, out: dict) -> None:
    import torch
    from gradscf_amd import molecule as M
    from gradscf_amd.basis import prepare_basis
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.packing import pack
    from gradscf_amd.scf import rhf

    for name, spec in (("C20H42", M.linear_alkane(20)), ("(H2O)32", M.water_cluster(32, seed=32, box=15.0))):
        atm, bas, env = pack(*prepare_basis(spec, "6-31g*", cart=True))
        plan = DirectJKPlan(atm, bas, env, cart=True, cutoff=1e-12)
        ref = rhf(plan, spec.nelectron, spec.nuclear_repulsion(), conv_tol=1e-10)
        mix = rhf(plan, spec.nelectron, spec.nuclear_repulsion(), conv_tol=1e-10, jk_flags={"mixed_threshold": thr})
        d = ref.density.contiguous()[None]

        def timed(**kw):
            plan.jk(d, symmetric=True, validate=False, **kw)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(10):
                plan.jk(d, symmetric=True, validate=False, **kw)
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / 10 * 1e3

        ms64, msmx = timed(), timed(mixed_threshold=thr)
        log(f"{name} 6-31G*: n={plan.nao} E={ref.energy:.10f} mixed dE={mix.energy - ref.energy:.2e} "
            f"J/K {ms64:.2f} ms FP64 / {msmx:.2f} ms mixed; SCF {ref.timings['scf_s']:.3f} / "
            f"{mix.timings['scf_s']:.3f} s")
        out[name] = {"n_ao": int(plan.nao), "quartets": int(plan.info["nquartets"]), "energy_fp64": ref.energy,
                     "energy_mixed": mix.energy, "energy_err": mix.energy - ref.energy,
                     "converged": bool(ref.converged and mix.converged), "jk_ms_fp64": ms64, "jk_ms_mixed": msmx,
                     "scf_s_fp64": ref.timings["scf_s"], "scf_s_mixed": mix.timings["scf_s"]}
    out["threshold"] = thr


def gradient_section(thr: float) -> dict:
    """(H2O)16 cc-pVDZ spherical: RHF nuclear gradient (all 48 nuclei) and
    dE/d(exponent, raw coefficient) for every primitive, in FP64 and from the
    mixed-precision SCF, with central finite-difference spot checks (1e-4 bohr,
    1e-5 relative) on the FP64 gradient."""
    import numpy as np
    import torch
    from gradscf_amd import molecule as M
    from gradscf_amd.basis import BasisParameters, prepare_basis
    from gradscf_amd.grad import rhf_basis_gradient, rhf_gradient, shell_nuclei
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.packing import pack
    from gradscf_amd.scf import rhf

    spec = M.water_cluster(16, seed=16, box=12.0)
    topo, par = prepare_basis(spec, "cc-pvdz", cart=False)
    atm, bas, env = pack(topo, par)
    sh = shell_nuclei(par.centers, par.nuclear_coords)
    nocc = spec.nelectron // 2

    def scf(env_, mixed=None, tol=1e-10):
        p = DirectJKPlan(atm, bas, env_, cart=False, cutoff=1e-12)
        flags = {} if mixed is None else {"mixed_threshold": mixed}
        return rhf(p, spec.nelectron, 0.0, conv_tol=tol, conv_tol_density=tol * 100, conv_tol_grad=tol * 10,
                   jk_flags=flags)

    res = scf(env, tol=1e-12)
    mix = scf(env, mixed=thr, tol=1e-12)
    gplan = DirectJKPlan(atm, bas, env, cart=False, cutoff=0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = rhf_gradient(gplan, res.density, res.mo_coeff, res.mo_energy, nocc, spec.charges, spec.coords_bohr, sh)
    torch.cuda.synchronize()
    t_nuc = time.perf_counter() - t0
    t0 = time.perf_counter()
    dexp, dcoef = rhf_basis_gradient(gplan, res.density, res.mo_coeff, res.mo_energy, nocc, bas, topo, par)
    torch.cuda.synchronize()
    t_par = time.perf_counter() - t0
    gm = rhf_gradient(gplan, mix.density, mix.mo_coeff, mix.mo_energy, nocc, spec.charges, spec.coords_bohr, sh)
    dexp_m, dcoef_m = rhf_basis_gradient(gplan, mix.density, mix.mo_coeff, mix.mo_energy, nocc, bas, topo, par)
    checks = []
    h = 1e-4
    for a, ax in ((0, 0), (17, 2)):
        rows = [a] + [len(spec.charges) + i for i in np.nonzero(sh == a)[0]]
        e = []
        for sgn in (1, -1):
            env_ = np.array(env, copy=True)
            for r in rows:
                env_[int(atm[r, 1]) + ax] += sgn * h
            e.append(scf(env_, tol=1e-12).energy)
        cp, cm = spec.coords_bohr.copy(), spec.coords_bohr.copy()
        cp[a, ax] += h
        cm[a, ax] -= h
        z = spec.charges.astype(float)

        def enuc(r):
            return sum(z[i] * z[j] / np.linalg.norm(r[i] - r[j]) for i in range(len(z)) for j in range(i))

        fd = (e[0] - e[1]) / (2 * h) + (enuc(cp) - enuc(cm)) / (2 * h)
        checks.append({"param": f"R[{a},{'xyz'[ax]}]", "analytic": float(g[a, ax]), "fd": fd,
                       "abs_err": abs(fd - float(g[a, ax]))})
    for s, p in ((2, 0), (4, 0)):
        hh = 1e-5 * par.exponents[s][p]
        e = []
        for sgn in (1, -1):
            ex = list(par.exponents)
            arr = np.array(ex[s], dtype=float)
            arr[p] += sgn * hh
            ex[s] = arr
            _, _, env_ = pack(topo, BasisParameters(tuple(ex), par.coefficients, par.centers, par.nuclear_coords))
            e.append(scf(env_, tol=1e-12).energy)
        fd = (e[0] - e[1]) / (2 * hh)
        checks.append({"param": f"alpha[shell {s}, prim {p}]", "analytic": float(dexp[s][p]), "fd": fd,
                       "abs_err": abs(fd - float(dexp[s][p]))})
    max_fd = max(c["abs_err"] for c in checks)
    dm = max(float(np.max(np.abs(gm - g))),
             max(float(np.max(np.abs(np.asarray(x) - np.asarray(y)))) for x, y in zip(dexp_m, dexp)),
             max(float(np.max(np.abs(np.asarray(x) - np.asarray(y)))) for x, y in zip(dcoef_m, dcoef)))
    log(f"gradient (H2O)16 cc-pVDZ: nuclear {t_nuc:.3f} s, basis parameters {t_par:.3f} s, "
        f"max |FD - analytic| {max_fd:.2e}, mixed vs FP64 {dm:.2e}")
    nexp = sum(len(a) for a in par.exponents)
    ncoef = sum(np.asarray(c).size for c in par.coefficients)
    return {"workload": "(H2O)16 cc-pVDZ spherical RHF; nuclear gradient of 48 nuclei and dE/d(exponent, "
                        "raw coefficient) of every primitive",
            "n_exponents": int(nexp), "n_coefficients": int(ncoef),
            "nuclear_gradient_s": t_nuc, "basis_gradient_s": t_par,
            "max_abs_nuclear_gradient": float(np.max(np.abs(g))),
            "translational_sum": float(np.max(np.abs(g.sum(axis=0)))),
            "fd_checks": checks, "max_fd_abs_err": max_fd, "mixed_threshold": thr,
            "mixed_vs_fp64_max_abs_gradient_diff": dm,
            "scf_energy_fp64": res.energy, "scf_energy_mixed": mix.energy}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-builds", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-scf", action="store_true", help="skip the CPU reference RHF of the workload")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision builds and SCF")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-molecule batched workload")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default per-class schedules")
    ap.add_argument("--no-incore", action="store_true",
                    help="skip the in-core integral builds and SCF (blocks stored once, digested per build)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the nonsymmetric-density and deterministic-mode builds (profiling runs: "
                         "every class kernel dispatch then belongs to the headline builds)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the water/6-31G, mixed C20H42/(H2O)32 and gradient sections")
    ap.add_argument("--retune", action="store_true",
                    help="run the autotuner even where a committed schedule table exists")
    ap.add_argument("--save-tuning", metavar="DIR", help="write autotuned schedule tables into DIR")
    ap.add_argument("--mixed-threshold", type=float, default=1e-7)
    ap.add_argument("--streams", type=int, default=None,
                    help="HIP streams the class launches of a build are spread over (engine default 4; "
                         "1 serialises them, as for a kernel trace whose per-kernel durations are compared "
                         "with the per-class timings)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the copy kernel that calibrates the HBM byte counters")
    ap.add_argument("--summarize-profile", nargs="+", metavar="DIR",
                    help="print a text summary of rocprofv3 CSV output directories and exit")
    ap.add_argument("--traffic-table", nargs=4, metavar=("FETCH", "WRITE", "CAL_FETCH", "CAL_WRITE"),
                    help="HBM bytes per class-kernel launch from rocprofv3 counter CSV directories")
    args = ap.parse_args()
    if args.summarize_profile:
        return summarize_profile(args.summarize_profile)
    if args.traffic_table:
        return traffic_table(*args.traffic_table)
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the J/K engine runs on a single MI355X "
              "(multi-GPU sharding is a design note, not an implemented path).", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from gradscf_amd import molecule as M
    from gradscf_amd import roofline
    from gradscf_amd.basis import prepare_basis
    from gradscf_amd.jk import DirectJKPlan
    from gradscf_amd.packing import pack
    from gradscf_amd.scf import rhf

    cutoff = 1e-12
    spec = M.water_cluster(16, seed=16, box=12.0)
    topo, par = prepare_basis(spec, "cc-pvdz", cart=False)
    atm, bas, env = pack(topo, par)
    log(f"workload (H2O)16 cc-pVDZ: {topo.nshell} shells, {topo.nao()} spherical AO")

    from gradscf_amd import autotune as at
    table_path = os.path.join(TUNING_DIR, "water16_ccpvdz_sph.json")
    have_table = not args.no_autotune and os.path.isfile(table_path) and not args.retune
    if have_table:
        at.load_table(table_path)   # every plan below uses the committed schedules
    else:
        at.reset()
    if args.streams is not None:
        at.set_options(streams=args.streams)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = DirectJKPlan(atm, bas, env, cart=False, cutoff=cutoff)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    log(f"plan: {plan.info['npairs']} segmented pairs, {plan.info['nquartets']} quartets, {plan_s * 1e3:.1f} ms")

    # the first SCF of the process also pays one-time start-up (GPU code
    # objects of the dense kernels, host LAPACK threads): it is reported as
    # wall_s_first, the steady-state SCF (same plan, run again) as wall_s
    scf_first = rhf(plan, spec.nelectron, spec.nuclear_repulsion(), conv_tol=1e-10)
    scf = rhf(plan, spec.nelectron, spec.nuclear_repulsion(), conv_tol=1e-10)
    dens = scf.density.contiguous()
    tuning = None
    if have_table:
        with open(table_path) as f:
            tuning = {"source": os.path.relpath(table_path, ROOT), "table": json.load(f)}
    elif not args.no_autotune:
        t0 = time.perf_counter()
        # a retune starts from the committed table (it may only improve on it)
        start = at.start_from_table(table_path) if os.path.isfile(table_path) else None
        result, opts = at.autotune_with_options(lambda: DirectJKPlan(atm, bas, env, cart=False, cutoff=cutoff),
                                                lambda p: p.jk(dens[None], symmetric=True, validate=False),
                                                reps=3, log=log, start=start)
        tuning = {"source": "autotuned in this run", "classes": result, "options": opts}
        if args.save_tuning:
            at.save_table(os.path.join(args.save_tuning, "water16_ccpvdz_sph.json"), result,
                          {"general_contraction": opts["general_contraction"]})
        tune_s = time.perf_counter() - t0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan = DirectJKPlan(atm, bas, env, cart=False, cutoff=cutoff)
        torch.cuda.synchronize()
        plan_s = time.perf_counter() - t0
        log(f"autotune {tune_s:.1f} s; tuned plan {plan_s * 1e3:.1f} ms")
        scf = rhf(plan, spec.nelectron, spec.nuclear_repulsion(), conv_tol=1e-10)
        dens = scf.density.contiguous()
    log(f"RHF E = {scf.energy:.10f} converged={scf.converged} cycles={scf.cycles} "
        f"scf {scf.timings['scf_s']:.3f} s (first in process {scf_first.timings['scf_s']:.3f} s; "
        f"{scf.timings['jk_builds']} J/K builds)")

    out = torch.empty((2, 1, plan.nao, plan.nao), dtype=torch.float64, device="cuda")
    for _ in range(args.warmup):
        plan.jk(dens[None], symmetric=True, validate=False, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.jk(dens[None], symmetric=True, validate=False, out=out)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = args.steps / elapsed
    log(f"J/K: {ms_per_step:.3f} ms/build, {value:.2f} builds/s")

    nonsym_ms, det_rec = None, None
    if not args.no_variants:
        # a nonsymmetric density (K needs D and D^T): the dd classes' exchange rows
        # of both exceed 64 KB together and run one density per launch
        dns = (dens + 1e-3 * torch.triu(dens, 1)).contiguous()
        out_ns = torch.empty_like(out)
        for _ in range(args.warmup):
            plan.jk(dns[None], symmetric=False, validate=False, out=out_ns)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.jk(dns[None], symmetric=False, validate=False, out=out_ns)
        torch.cuda.synchronize()
        nonsym_ms = (time.perf_counter() - t0) / args.steps * 1e3
        log(f"nonsymmetric density: {nonsym_ms:.3f} ms/build")

        # deterministic mode (integer accumulation, bitwise repeatable) on the same workload
        before_opts = at.get_options()
        at.set_options(deterministic=1)
        plan_det = DirectJKPlan(atm, bas, env, cart=False, cutoff=cutoff)
        at.set_options(**before_opts)
        out_det = torch.empty_like(out)
        for _ in range(args.warmup):
            plan_det.jk(dens[None], symmetric=True, validate=False, out=out_det)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan_det.jk(dens[None], symmetric=True, validate=False, out=out_det)
        torch.cuda.synchronize()
        det_ms = (time.perf_counter() - t0) / args.steps * 1e3
        det_again = plan_det.jk(dens[None], symmetric=True, validate=False)
        det_rec = {"ms_per_step": det_ms, "bitwise_repeatable": bool(torch.equal(det_again, out_det)),
                   "max_abs_jk_diff_vs_fp64_atomics": float((out_det - out).abs().max())}
        del plan_det
        log(f"deterministic mode: {det_ms:.3f} ms/build, repeatable {det_rec['bitwise_repeatable']}, "
            f"max|dJK| vs atomics {det_rec['max_abs_jk_diff_vs_fp64_atomics']:.1e}")

    # mixed precision: same workload, FP32 for quartets with Q_ij Q_kl max|D| < threshold
    thr = args.mixed_threshold
    mixed = None
    if not args.no_mixed:
        out_mixed = torch.empty_like(out)
        for _ in range(args.warmup):
            plan.jk(dens[None], symmetric=True, validate=False, out=out_mixed, mixed_threshold=thr)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.jk(dens[None], symmetric=True, validate=False, out=out_mixed, mixed_threshold=thr)
        torch.cuda.synchronize()
        mixed_ms = (time.perf_counter() - t0) / args.steps * 1e3
        mixed_err = float((out_mixed - out).abs().max())
        scf_mixed = rhf(plan, spec.nelectron, spec.nuclear_repulsion(), conv_tol=1e-10,
                        jk_flags={"mixed_threshold": thr})
        log(f"mixed (thr {thr:g}): {mixed_ms:.3f} ms/build, max|dJK| {mixed_err:.2e}, "
            f"E {scf_mixed.energy:.10f} (dE {scf_mixed.energy - scf.energy:.2e}), "
            f"scf {scf_mixed.timings['scf_s']:.3f} s")
        mixed = {"threshold": thr, "ms_per_step": mixed_ms, "builds_per_s": 1e3 / mixed_ms,
                 "speedup_vs_fp64": ms_per_step / mixed_ms, "max_abs_jk_diff_vs_fp64": mixed_err,
                 "scf_energy": scf_mixed.energy, "scf_energy_err_vs_fp64": scf_mixed.energy - scf.energy,
                 "scf_wall_s": scf_mixed.timings["scf_s"], "scf_converged": bool(scf_mixed.converged),
                 "pays": ms_per_step / mixed_ms > 1.0,
                 "note": "non-packed FP32 VALU issues at the FP64 rate on CDNA4 (packed FP32 is excluded "
                         "from this build), so FP32 quartets do not shorten a build; the batched "
                         "speed-up over per-molecule FP64 calls comes from batching"}

    # per-class kernel durations (HIP events on the launch stream) over a timed pass
    plan.profile(True)
    for _ in range(args.steps):
        plan.jk(dens[None], symmetric=True, validate=False, out=out)
    torch.cuda.synchronize()
    times = plan.class_times()
    plan.profile(False)
    work = roofline.build_work(bas, plan.schwarz(), cutoff, cart=False)
    per_class = []
    for cls, ms, launches in times:
        w = work["classes"].get(cls)
        avg = ms / max(launches, 1)
        rec = {"class": "".join("spdf"[v] for v in cls), "avg_ms": avg}
        if w:
            rec["tflops"] = w["flop_eval"] / (avg * 1e-3) / 1e12
            rec["flop_eval"] = w["flop_eval"]
            # algorithmic units of one launch: surviving canonical quartets
            # (parent shells) and their primitive quartets
            rec["quartets"] = w["quartets"]
            rec["prim_quartets"] = w["prim_quartets"]
        per_class.append(rec)
    per_class.sort(key=lambda r: -r["avg_ms"])
    dom = per_class[0]

    # in-core integrals (the packed-ERI route): every screened-in quartet block
    # stored once in HBM, later builds digest stored blocks; reported beside
    # the direct builds, never as `value`
    incore = None
    if not args.no_incore:
        nbytes = plan.incore_bytes()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.incore(True)
        torch.cuda.synchronize()
        store_ms = (time.perf_counter() - t0) * 1e3
        out_in = torch.empty_like(out)
        for _ in range(args.warmup):
            plan.jk(dens[None], symmetric=True, validate=False, out=out_in)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.jk(dens[None], symmetric=True, validate=False, out=out_in)
        torch.cuda.synchronize()
        in_ms = (time.perf_counter() - t0) / args.steps * 1e3
        in_err = float((out_in - out).abs().max())
        scf_in = rhf(plan, spec.nelectron, spec.nuclear_repulsion(), conv_tol=1e-10)
        plan.incore(False)
        log(f"in-core: {nbytes / 1e9:.2f} GB stored in {store_ms:.1f} ms, {in_ms:.3f} ms/build "
            f"({ms_per_step / in_ms:.2f}x direct), max|dJK| {in_err:.2e}, SCF {scf_in.timings['scf_s']:.3f} s "
            f"(dE {scf_in.energy - scf.energy:.2e})")
        incore = {"bytes": nbytes, "store_ms": store_ms, "ms_per_step": in_ms, "builds_per_s": 1e3 / in_ms,
                  "speedup_vs_direct": ms_per_step / in_ms, "max_abs_jk_diff_vs_direct": in_err,
                  "scf_wall_s": scf_in.timings["scf_s"], "scf_wall_s_including_store": scf_in.timings["scf_s"]
                  + store_ms * 1e-3, "scf_energy_err_vs_direct": scf_in.energy - scf.energy,
                  "scf_converged": bool(scf_in.converged)}
    total_kernel_ms = sum(r["avg_ms"] for r in per_class)
    tot = work["total"]
    achieved = dom.get("tflops", 0.0)
    j_gpu = out[0, 0].cpu().numpy()
    k_gpu = out[1, 0].cpu().numpy()

    result = {
        "metric": "J/K builds/s + SCF wall time, mixed vs FP64 on 1 MI355X at <=1 uEh error",
        "value": value,
        "unit": "J/K builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "(H2O)16 cc-pVDZ spherical, plain FP64 direct J/K at the converged RHF density",
                   "n_ao": int(plan.nao), "shells": int(topo.nshell), "schwarz_cutoff": cutoff,
                   "geometry": "seeded water cluster: seed 16, 12 A cube, O-O >= 2.6 A",
                   "parallelism": "single GPU"},
        "scf": {"energy": scf.energy, "converged": bool(scf.converged), "cycles": scf.cycles,
                "wall_s": scf.timings["scf_s"], "wall_s_first": scf_first.timings["scf_s"],
                "integrals_s": scf.timings["integrals_s"], "jk_builds": scf.timings["jk_builds"]},
        "mixed": mixed,
        "nonsymmetric_ms_per_step": nonsym_ms,
        "deterministic": det_rec,
        "incore": incore,
        "plan_ms": plan_s * 1e3,
        "ms_per_step_including_plan": ms_per_step + plan_s * 1e3,
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": profiled_traffic(dom["class"]),
                     "kernel": f"jk class ({dom['class']}) quartet kernel",
                     "note": "FP64 vector-ALU Rys quartet evaluation (no MFMA shape); algorithmic FLOP_eval "
                             "of the class / avg HIP-event duration; traffic = HBM bytes per launch from the "
                             "committed FETCH_SIZE/WRITE_SIZE summary"},
        "build_roofline": {"flop_eval": tot["flop_eval"], "flop_dig": tot["flop_dig"],
                           "bytes_dig": tot["bytes_dig"], "quartets": tot["quartets"],
                           "class_kernels_ms": total_kernel_ms,
                           "tflops_eval": tot["flop_eval"] / (total_kernel_ms * 1e-3) / 1e12},
        "classes": per_class,
        "autotune": tuning,
        "engine_options": at.get_options(),
        "measured_peaks": measured_peaks(),
    }
    mp = result["measured_peaks"]
    log(f"measured peaks: FP64 FMA {mp['fp64_fma_tflops']:.1f} TFLOP/s, device copy {mp['copy_gbs']:.0f} GB/s")
    water = None
    if not args.no_extra:
        water = water_section()
        result["water_631g"] = water["record"]
        # (the config-3 table replaces the headline one only when that came from its file)
        result["mixed_configs"] = mixed_section(thr, tuned=have_table)
        result["gradients"] = gradient_section(thr)
    result["batch"] = None if args.no_batch else batch_section(thr, autotune=not args.no_autotune,
                                                               retune=args.retune, save_dir=args.save_tuning,
                                                               incore=not args.no_incore)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(atm, bas, env, dens.cpu().numpy(), cutoff, args.cpu_builds,
                                              j_gpu, k_gpu, None if args.no_cpu_scf else (spec, scf.energy))
        if water is not None:
            result["cpu_baseline"]["water_631g"] = cpu_baseline_water(water)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
