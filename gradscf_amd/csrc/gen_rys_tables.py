"""Generate the Rys-quadrature interpolation table compiled into the HIP library.

For n = 1..NMAX roots of the weight exp(-x t^2) on t in [0, 1] (written in
s = t^2), the roots s_i(x) and weights w_i(x) are tabulated on x in [0, XMAX)
as degree-DEG polynomials in y = (x - centre)/(H/2) on intervals of width H
(a Chebyshev fit at the Chebyshev nodes, converted to monomial coefficients in
60-digit arithmetic for Horner evaluation: one FMA per degree, with the same
accuracy as Clenshaw's recurrence since the coefficients decay); above XMAX the
large-x limit (positive Gauss-Hermite nodes r_i of order 2n: s_i = r_i^2/x,
w_i = W_i/sqrt(x)) is exact to double precision.

For one and two roots the kernels instead use Boys functions F_0..F_3 from
a K-term Taylor expansion about the nearest point of a grid T_i = i H_B
(dF_m/dT = -F_{m+1}, so F_m(T_i - d) = sum_k F_{m+k}(T_i) d^k / k!, |d| <=
H_B/2) and the closed-form one- and two-point quadratures of those moments;
this module also tabulates F_0..F_{NB-1} on that grid (BOYS_* below).

Reference values are computed in 60-digit arithmetic (mpmath): Boys moments
F_k(x) = 1F1(k+1/2; k+3/2; -x)/(2k+1), the Chebyshev algorithm for the
recurrence coefficients of the measure, and the eigen-decomposition of the
Jacobi matrix.  Output: csrc/_generated/rys_table.inc (constants) and rys_table_body.inc
(the coefficient list, included into a __device__ array).

Usage: python gen_rys_tables.py [--out PATH]
"""
from __future__ import annotations

import argparse
import hashlib
import os
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

NMAX = 7
XMAX = 80.0
H = 2.0
DEG = 13
DPS = 60
# Boys grid: exact binary spacing, Taylor terms, functions per grid point
# (F_0..F_3 need F_0..F_{3+K-1}; padded to a 16-byte multiple)
BOYS_H = 0.0625
BOYS_TMAX = 48.0
BOYS_K = 8
BOYS_NB = 12


def _rys_mp(n: int, x: float):
    import mpmath as mp
    mp.mp.dps = DPS
    x = mp.mpf(x)
    mu = [mp.hyp1f1(k + mp.mpf(1) / 2, k + mp.mpf(3) / 2, -x) / (2 * k + 1) for k in range(2 * n)]
    # Chebyshev algorithm: ordinary moments -> recurrence coefficients
    alpha = [mp.mpf(0)] * n
    beta = [mp.mpf(0)] * n
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = list(mu)
    alpha[0] = mu[1] / mu[0]
    beta[0] = mu[0]
    for k in range(1, n):
        new = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha[k] = new[k + 1] / new[k] - sig[k] / sig[k - 1]
        beta[k] = new[k] / sig[k - 1]
        sig_prev, sig = sig, new
    J = mp.zeros(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    E, Q = mp.eigsy(J)
    order = sorted(range(n), key=lambda i: E[i])
    roots = [E[i] for i in order]
    weights = [beta[0] * Q[0, i] ** 2 for i in order]
    return [float(r) for r in roots], [float(w) for w in weights]


def _interval_coeffs(args):
    n, k = args
    import mpmath as mp
    mp.mp.dps = DPS
    a = k * H
    nodes = [mp.cos(mp.pi * (j + mp.mpf(1) / 2) / (DEG + 1)) for j in range(DEG + 1)]
    vals = []
    for y in nodes:
        r, w = _rys_mp(n, float(a + H / 2 + (H / 2) * y))
        vals.append(r + w)
    # Chebyshev coefficients c_m = 2/(D+1) sum_j f(y_j) T_m(y_j) (c_0 halved)
    cheb = []
    for m in range(DEG + 1):
        row = []
        for f in range(2 * n):
            s = mp.mpf(0)
            for j, y in enumerate(nodes):
                s += vals[j][f] * mp.cos(m * mp.acos(y))
            s *= mp.mpf(2) / (DEG + 1)
            if m == 0:
                s /= 2
            row.append(s)
        cheb.append(row)
    # monomial coefficients a_k = sum_m c_m [y^k] T_m(y), T_m from T_m = 2y T_{m-1} - T_{m-2}
    tm = [[0] * (DEG + 1) for _ in range(DEG + 1)]
    tm[0][0] = 1
    tm[1][1] = 1
    for m in range(2, DEG + 1):
        for kk in range(DEG + 1):
            tm[m][kk] = (2 * tm[m - 1][kk - 1] if kk else 0) - tm[m - 2][kk]
    coeffs = []
    for kk in range(DEG + 1):
        coeffs.append([float(sum(cheb[m][f] * tm[m][kk] for m in range(DEG + 1))) for f in range(2 * n)])
    return n, k, coeffs


def _boys_row(i: int):
    """F_0..F_{NB-1} at T_i = i * BOYS_H (the double the device forms)."""
    import mpmath as mp
    mp.mp.dps = DPS
    t = mp.mpf(float(i * BOYS_H))
    return i, [float(mp.hyp1f1(m + mp.mpf(1) / 2, m + mp.mpf(3) / 2, -t) / (2 * m + 1)) for m in range(BOYS_NB)]


def _hermite(n: int):
    import numpy as np
    from numpy.polynomial.hermite import hermgauss
    x, w = hermgauss(2 * n)
    pos = x > 0
    r, W = x[pos], w[pos]
    o = np.argsort(r)
    return [float(v) for v in r[o] ** 2], [float(v) for v in W[o]]


def generate():
    nint = int(XMAX / H)
    jobs = [(n, k) for n in range(1, NMAX + 1) for k in range(nint)]
    npt = int(BOYS_TMAX / BOYS_H) + 1
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = list(ex.map(_interval_coeffs, jobs, chunksize=4))
        boys = dict(ex.map(_boys_row, range(npt), chunksize=16))
    table = {(n, k): c for n, k, c in results}
    lines = ["// Generated by gen_rys_tables.py -- do not edit.",
             f"#define RYS_NMAX {NMAX}", f"#define RYS_XMAX {XMAX!r}", f"#define RYS_H {H!r}",
             f"#define RYS_DEG {DEG}", f"#define RYS_NINT {nint}"]
    offsets = [0]
    flat = []
    for n in range(1, NMAX + 1):
        for k in range(nint):
            for row in table[(n, k)]:
                flat.extend(row)
        offsets.append(len(flat))
    lines.append("constexpr int rys_table_offset[%d] = {%s};" % (
        len(offsets), ", ".join(str(o) for o in offsets)))
    lines.append(f"#define RYS_TABLE_SIZE {len(flat)}")
    body = [", ".join(float(v).hex() for v in flat[i:i + 4]) + "," for i in range(0, len(flat), 4)]
    hr, hw = [], []
    hoff = [0]
    for n in range(1, NMAX + 1):
        r, w = _hermite(n)
        hr.extend(r)
        hw.extend(w)
        hoff.append(len(hr))
    lines.append("constexpr double rys_asym_root_c[%d] = {%s};" % (len(hr), ", ".join(v.hex() for v in hr)))
    lines.append("constexpr double rys_asym_weight_c[%d] = {%s};" % (len(hw), ", ".join(v.hex() for v in hw)))
    lines += [f"#define BOYS_H {BOYS_H!r}", f"#define BOYS_TMAX {BOYS_TMAX!r}", f"#define BOYS_K {BOYS_K}",
              f"#define BOYS_NB {BOYS_NB}", f"#define BOYS_NPT {npt}"]
    bflat = [v for i in range(npt) for v in boys[i]]
    bbody = [", ".join(float(v).hex() for v in bflat[i:i + 4]) + "," for i in range(0, len(bflat), 4)]
    return "\n".join(lines) + "\n", "\n".join(body) + "\n", "\n".join(bbody) + "\n"


def source_hash() -> str:
    return hashlib.sha256(Path(__file__).read_bytes()).hexdigest()[:16]


def ensure(out: Path) -> Path:
    """(Re)generate `out` (+ the table body beside it) unless it carries this generator's hash."""
    tag = f"// source-hash {source_hash()}\n"
    body_path = out.with_name("rys_table_body.inc")
    boys_path = out.with_name("boys_table_body.inc")
    if out.is_file() and body_path.is_file() and boys_path.is_file() and out.read_text().startswith(tag):
        return out
    out.parent.mkdir(parents=True, exist_ok=True)
    head, body, bbody = generate()
    for path, text in ((body_path, body), (boys_path, bbody), (out, tag + head)):
        tmp = path.with_suffix(".tmp")
        tmp.write_text(text)
        tmp.replace(path)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", type=Path, default=Path(__file__).with_name("_generated") / "rys_table.inc")
    print(ensure(ap.parse_args().out))
