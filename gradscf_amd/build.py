"""Build the HIP engine in-tree: gradscf_amd/lib/libgradscf_jk.so (gfx950).

Steps: generate the Rys table (csrc/gen_rys_tables.py, cached by source hash),
compile one translation unit per angular-momentum class (and per auxiliary
angular momentum of the density-fitting kernels) plus the host engine
with hipcc in parallel, and link with an export map that keeps only gscf_*.

Usage: python -m gradscf_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "lib" / "libgradscf_jk.so"

CLASSES = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 0), (1, 1, 1, 1),
           (2, 0, 0, 0), (2, 0, 1, 0), (2, 0, 1, 1), (2, 0, 2, 0), (2, 1, 0, 0), (2, 1, 1, 0),
           (2, 1, 1, 1), (2, 1, 2, 0), (2, 1, 2, 1), (2, 2, 0, 0), (2, 2, 1, 0), (2, 2, 1, 1),
           (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2)]

# auxiliary shells of the density-fitting integrals: up to g
DF_MAX_L = 4

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + str(ROOT / "include"), "-I" + str(CSRC)]


def _compile(src: Path, obj: Path, defines=()) -> None:
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name} {defines}:\n{r.stderr[-4000:]}")


def build(jobs: int | None = None) -> Path:
    # the generator fans out over processes, so it runs as its own program
    # (it regenerates only when its source hash changed)
    r = subprocess.run([sys.executable, str(CSRC / "gen_rys_tables.py"), "--out",
                        str(CSRC / "_generated" / "rys_table.inc")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Rys table generation failed:\n{r.stderr[-4000:]}")
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    units = [(CSRC / "jk_engine.hip", BUILD / "jk_engine.o", ()),
             (CSRC / "int1e.hip", BUILD / "int1e.o", ())]
    for c in CLASSES:
        tag = "".join(map(str, c))
        units.append((CSRC / "jk_class.hip", BUILD / f"jk_class_{tag}.o",
                      (f"GSCF_LA={c[0]}", f"GSCF_LB={c[1]}", f"GSCF_LC={c[2]}", f"GSCF_LD={c[3]}")))
    for lc in range(DF_MAX_L + 1):
        units.append((CSRC / "df_class.hip", BUILD / f"df_class_{lc}.o", (f"GSCF_DF_LC={lc}",)))
    jobs = jobs or min(16, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda u: _compile(*u), units))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", str(tmp), *[str(u[1]) for u in units],
           "-Wl,--version-script=" + str(CSRC / "exports.map")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    print(build(ap.parse_args().jobs))
