"""Autotuner for the per-class J/K schedules.

Each canonical quartet class runs as its own kernel launch, so its schedule
(bra-chunk size, bras per chunk, LDS packing, Rys-table placement, primitive
split) can be chosen independently of the other classes.  The tuner times
every class under candidate schedules with HIP events on the launch stream
(coordinate descent from the defaults, one parameter at a time), keeps the
fastest schedule per class and installs the result in the engine's
process-wide tuning table, which plans copy when they are created.

Results never change J/K beyond the order of FP64 additions (tested).
"""
from __future__ import annotations

import ctypes
from dataclasses import asdict, dataclass

from . import _native

PAIR_CLASSES = ((0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2))
LETTERS = "spdf"


@dataclass(frozen=True)
class Schedule:
    chunk_items: int = 4096
    bras_max: int = 16
    lds_pack: int = 32 * 1024
    rys_lds: int = 0
    target_prim: int = 48
    threads: int = 256
    ket_lds: int = 0


# Coordinate descent, one entry at a time; a tuple key is searched jointly
# (chunk size and primitive split interact: both set the items per chunk).
SEARCH = {
    "rys_lds": (0, 1),
    "ket_lds": (0, 1),
    "threads": (256, 512),
    ("chunk_items", "target_prim"): tuple((c, t) for c in (512, 1024, 2048, 4096, 8192, 16384)
                                          for t in (6, 12, 24, 48, 96, 192)),
    "bras_max": (1, 2, 4, 8, 16, 32, 64),
    "lds_pack": (4096, 8192, 16384, 32768, 49152, 65536),
}


def class_index(cls) -> int:
    """(la, lb, lc, ld) -> index of the engine's tuning table."""
    cb = PAIR_CLASSES.index((cls[0], cls[1]))
    ck = PAIR_CLASSES.index((cls[2], cls[3]))
    return cb * (cb + 1) // 2 + ck


def class_name(cls) -> str:
    return "".join(LETTERS[v] for v in cls)


def set_schedule(schedule: Schedule, cls=None) -> None:
    t = _native.ClassTuning(**asdict(schedule))
    _native.check(_native.lib().gscf_tuning_set(-1 if cls is None else class_index(cls), ctypes.byref(t)))


def get_schedule(cls) -> Schedule:
    t = _native.ClassTuning()
    _native.check(_native.lib().gscf_tuning_get(class_index(cls), ctypes.byref(t)))
    return Schedule(*(getattr(t, f) for f, _ in _native.ClassTuning._fields_))


def reset() -> None:
    _native.check(_native.lib().gscf_tuning_reset())


def get_options() -> dict:
    """Process-wide engine options (general contraction, launch streams,
    partial-sum flushes)."""
    o = _native.Options()
    _native.check(_native.lib().gscf_options_get(ctypes.byref(o)))
    return {f: getattr(o, f) for f, _ in _native.Options._fields_}


def set_options(**kw) -> None:
    """Change engine options; plans created afterwards use them."""
    cur = get_options()
    unknown = set(kw) - set(cur)
    if unknown:
        raise ValueError(f"unknown engine options {sorted(unknown)}")
    cur.update(kw)
    o = _native.Options(**cur)
    _native.check(_native.lib().gscf_options_set(ctypes.byref(o)))


def build_ms(make_plan, run, reps: int = 5) -> float:
    """Average wall time of run(plan) on the current stream (HIP events), with
    class launches spread over streams as in production."""
    import torch
    plan = make_plan()
    run(plan)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        run(plan)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def choose_options(make_plan, run, *, reps: int = 5, log=None) -> dict:
    """Time whole builds with and without native general contraction and keep
    the faster; returns {"general_contraction": 0|1, "ms": {...}}."""
    ms = {}
    for gc in (1, 0):
        set_options(general_contraction=gc)
        ms[gc] = build_ms(make_plan, run, reps)
        if log:
            log(f"options general_contraction={gc}: {ms[gc]:.3f} ms per build")
    best = min(ms, key=ms.get)
    set_options(general_contraction=best)
    return {"general_contraction": best, "ms": ms}


def _class_times(make_plan, run, reps):
    import torch
    plan = make_plan()
    run(plan)  # warm-up (workspace, caches)
    torch.cuda.synchronize()
    plan.profile(True)
    for _ in range(reps):
        run(plan)
    torch.cuda.synchronize()
    times = {cls: ms / max(n, 1) for cls, ms, n in plan.class_times()}
    plan.profile(False)
    return times


def autotune(make_plan, run, *, reps: int = 3, search=None, passes: int = 2, log=None,
             confirm_tol: float = 0.15, start=None) -> dict:
    """Choose a schedule per class.

    make_plan(): builds a DirectJKPlan under the currently installed tuning.
    run(plan): one J/K build of the workload on that plan.
    passes: sweeps of the coordinate descent (the second revisits early
    choices, e.g. the root evaluation, under the later ones).
    A confirmation pass then times every class under the final table: a
    class that runs more than `confirm_tol` slower than it did when its
    schedule was chosen (a choice measured beside other classes' trial
    schedules) is timed again with each schedule it held earlier, newest
    first, and keeps the fastest it confirmed.
    start: optional {class: Schedule} (e.g. a committed table, see
    start_from_table) to descend from instead of the default schedule; the
    descent only accepts improvements, so a retune from a table never ends
    slower than the table (coordinate descent from the defaults can settle in
    a worse local optimum than a table found along another path).
    Returns {class name: {"schedule": dict, "ms": confirmed avg ms,
    "default_ms": ms at the start}} and leaves the chosen schedules installed."""
    search = SEARCH if search is None else search
    reset()
    base = Schedule()
    for cls, sch in (start or {}).items():
        set_schedule(sch, cls)
    best_t = _class_times(make_plan, run, reps)
    default_t = dict(best_t)
    best_s = {cls: (start or {}).get(cls, base) for cls in best_t}
    earlier = {cls: [] for cls in best_t}  # schedules a class held before its current one
    for param, values in [kv for _ in range(passes) for kv in search.items()]:
        names = param if isinstance(param, tuple) else (param,)
        current = {cls: best_s[cls] for cls in best_t}
        for v in values:
            upd = dict(zip(names, v if isinstance(param, tuple) else (v,)))
            if all(all(getattr(current[c], k) == x for k, x in upd.items()) for c in current):
                continue
            # every class tries this value on top of its current best
            for cls in current:
                set_schedule(Schedule(**{**asdict(current[cls]), **upd}), cls)
            t = _class_times(make_plan, run, reps)
            for cls, ms in t.items():
                if cls in best_t and ms < best_t[cls]:
                    best_t[cls] = ms
                    cand = Schedule(**{**asdict(current[cls]), **upd})
                    if cand != best_s[cls] and best_s[cls] not in earlier[cls]:
                        earlier[cls].append(best_s[cls])
                    best_s[cls] = cand
            if log:
                log(f"autotune {param}={v}: {sum(t.values()):.3f} ms over classes")
        for cls in best_s:
            set_schedule(best_s[cls], cls)
    for cls in best_s:
        set_schedule(best_s[cls], cls)
    final_t = _class_times(make_plan, run, reps)
    for cls in best_s:
        final_t.setdefault(cls, best_t[cls])
    suspect = [cls for cls in best_s if final_t[cls] > (1.0 + confirm_tol) * best_t[cls]]
    for depth in range(max((len(earlier[c]) for c in suspect), default=0)):
        trial = {cls: earlier[cls][-1 - depth] for cls in suspect if depth < len(earlier[cls])}
        if not trial:
            break
        for cls, sch in trial.items():
            set_schedule(sch, cls)
        t = _class_times(make_plan, run, reps)
        for cls, sch in trial.items():
            if t.get(cls, float("inf")) < final_t[cls]:
                final_t[cls] = t[cls]
                best_s[cls] = sch
            set_schedule(best_s[cls], cls)
        if log:
            log(f"autotune confirmation: {len(trial)} classes retried earlier schedules")
    for cls in best_s:
        set_schedule(best_s[cls], cls)
    return {class_name(cls): {"schedule": asdict(best_s[cls]), "ms": final_t[cls], "default_ms": default_t[cls]}
            for cls in best_s}


def start_from_table(path) -> dict:
    """{class: Schedule} of a table written by save_table (the start point
    of a retune that may only improve on it)."""
    import json
    with open(path) as f:
        table = json.load(f)
    return {tuple(LETTERS.index(ch) for ch in name): Schedule(**sch)
            for name, sch in table.items() if not name.startswith("_")}


def autotune_with_options(make_plan, run, *, reps: int = 3, log=None, start=None):
    """Tune the per-class schedules once with native general contraction and
    once without, keep whichever gives the faster whole build, and install it.
    Returns (per-class result as from autotune, {"general_contraction": ...,
    "build_ms": {0: ms, 1: ms}})."""
    runs = {}
    for gc in (1, 0):
        set_options(general_contraction=gc)
        res = autotune(make_plan, run, reps=reps, log=log, start=start)
        ms = build_ms(make_plan, run)
        runs[gc] = (ms, res)
        if log:
            log(f"autotune general_contraction={gc}: {ms:.3f} ms per build")
    gc = min(runs, key=lambda k: runs[k][0])
    set_options(general_contraction=gc)
    reset()
    for name, v in runs[gc][1].items():
        set_schedule(Schedule(**v["schedule"]), tuple(LETTERS.index(ch) for ch in name))
    return runs[gc][1], {"general_contraction": gc, "build_ms": {k: v[0] for k, v in runs.items()}}


def save_table(path, result: dict, options: dict | None = None) -> None:
    """Write the chosen per-class schedules ({class name: {"schedule": ...}})
    and, when given, the engine options under the key "_options"."""
    import json
    table = {k: v["schedule"] for k, v in sorted(result.items())}
    if options:
        table["_options"] = dict(options)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")


def load_table(path) -> dict:
    """Install per-class schedules (and engine options, if recorded) from a
    file written by save_table; classes not listed keep the defaults.
    Returns the table."""
    import json
    reset()
    with open(path) as f:
        table = json.load(f)
    for name, sched in table.items():
        if name == "_options":
            set_options(**sched)
            continue
        cls = tuple(LETTERS.index(ch) for ch in name)
        set_schedule(Schedule(**sched), cls)
    return table
