"""Algorithmic work of one direct J/K build (the frozen roofline formula).

Sums over the canonical parent-shell quartets q = (ij|kl) (i>=j, k>=l, ij>=kl)
that survive the reference's Schwarz test Q_ij Q_kl >= cutoff
(src/gradscf/integrals/_native/csrc/coulomb_shells.h:98-110):

  P_q  = np_i np_j np_k np_l           primitive quartets
  nf_q = prod ncart(l)                  Cartesian components
  nr_q = floor((l_i+l_j+l_k+l_l)/2)+1   Rys roots
  V_q  = (l_i+l_j+1)(l_k+l_l+1)         2D-integral entries
  nc_q = prod nctr                      contraction columns

  FLOP_eval = sum_q P_q [40 + 12 nr + nr (18 V + 3 nf)] + sum_q 2 nf P_q nc_q
  FLOP_dig  = sum_q 12 U_q,   U_q = unique AO quartets (a>=b, c>=d, ab>=cd) in q
  BYTES_dig = sum_q 24 (di dj + dk dl + di dk + di dl + dj dk + dj dl)

(40 = per-primitive-quartet prefactor/T/exp set-up; 12 nr = roots and weights;
18 V per root = two FMAs per 2D entry per axis; 3 nf per root = Ix Iy Iz w
accumulation; 2 nf nc = contraction FMA; d = AO block dimensions nctr x
(Cartesian or spherical) components.)  Because every term is symmetric under
exchanging the bra and ket pair, the canonical sum is evaluated as half the sum
over ordered pair-pairs plus half the diagonal pair-pairs, pair class by pair
class, with the Schwarz test applied through sorted bounds.
"""
from __future__ import annotations

import numpy as np


def _ncart(l):
    return (l + 1) * (l + 2) // 2


def pair_table(bas: np.ndarray, q: np.ndarray | None, cart: bool):
    """Per parent pair i>=j: l_i, l_j, P, nc, dims, Schwarz bound."""
    nbas = len(bas)
    ii, jj = np.tril_indices(nbas)
    l, npr, nct = bas[:, 1], bas[:, 2], bas[:, 3]
    comp = np.where(cart, (l + 1) * (l + 2) // 2, 2 * l + 1)
    d = nct * comp
    qv = np.ones(len(ii)) if q is None else q[ii, jj]
    la, lb = np.maximum(l[ii], l[jj]), np.minimum(l[ii], l[jj])
    return dict(i=ii, j=jj, la=la, lb=lb, l1=l[ii], l2=l[jj], P=npr[ii] * npr[jj], nc=nct[ii] * nct[jj],
                di=d[ii], dj=d[jj], same=(ii == jj), q=qv)


def build_work(bas: np.ndarray, q: np.ndarray | None, cutoff: float, cart: bool) -> dict:
    """Algorithmic FLOP/bytes per class {(la, lb, lc, ld): {...}} and totals."""
    bas = np.asarray(bas)
    pt = pair_table(bas, q, cart)
    keep = np.ones(len(pt["q"]), dtype=bool)
    if cutoff > 0:
        keep = pt["q"] * pt["q"].max() >= cutoff
    for k in pt:
        pt[k] = pt[k][keep]
    pcls = pt["la"] * (pt["la"] + 1) // 2 + pt["lb"]
    nab = np.where(pt["same"], pt["di"] * (pt["di"] + 1) // 2, pt["di"] * pt["dj"])
    out = {}
    classes = sorted(set(pcls.tolist()))
    for cb in classes:
        B = np.where(pcls == cb)[0]
        for ck in classes:
            if ck > cb:
                continue
            K = np.where(pcls == ck)[0]
            order = np.argsort(-pt["q"][K])
            K = K[order]
            qk = pt["q"][K]
            lb_ = (int(pt["la"][B[0]]), int(pt["lb"][B[0]]))
            lk_ = (int(pt["la"][K[0]]), int(pt["lb"][K[0]]))
            ltot = sum(lb_) + sum(lk_)
            nr = ltot // 2 + 1
            V = (sum(lb_) + 1) * (sum(lk_) + 1)
            nf = _ncart(lb_[0]) * _ncart(lb_[1]) * _ncart(lk_[0]) * _ncart(lk_[1])
            # prefix sums over Schwarz-sorted kets
            cs = lambda x: np.concatenate([[0.0], np.cumsum(x.astype(np.float64))])
            cP, cPnc = cs(pt["P"][K]), cs(pt["P"][K] * pt["nc"][K])
            cnab = cs(nab[K])
            dk, dl = pt["di"][K].astype(np.float64), pt["dj"][K].astype(np.float64)
            cdkdl, cdk, cdl = cs(dk * dl), cs(dk), cs(dl)
            if cutoff > 0:
                cnt = np.searchsorted(-qk, -(cutoff / pt["q"][B]), side="right")
            else:
                cnt = np.full(len(B), len(K))
            Pb, ncb = pt["P"][B].astype(np.float64), pt["nc"][B].astype(np.float64)
            di, dj = pt["di"][B].astype(np.float64), pt["dj"][B].astype(np.float64)
            sumP = np.sum(Pb * cP[cnt])                      # ordered sum of P_q
            sumPnc = np.sum(Pb * ncb * cPnc[cnt])
            sumU = np.sum(nab[B] * cnab[cnt])
            sumB = np.sum(24 * (di * dj * cnt + cdkdl[cnt] + (di + dj) * (cdk[cnt] + cdl[cnt])))
            nq = float(np.sum(cnt))
            # diagonal (bra == ket) terms exist only within one class
            if cb == ck:
                dq = pt["q"][B] * pt["q"][B] >= cutoff if cutoff > 0 else np.ones(len(B), bool)
                dP = np.sum((Pb * Pb)[dq])
                dPnc = np.sum((Pb * Pb * ncb * ncb)[dq])
                dU = np.sum((nab[B] * (nab[B] + 1) / 2.0)[dq])
                dUo = np.sum((nab[B].astype(np.float64) ** 2)[dq])
                dB = np.sum((24 * (2 * di * dj + (di + dj) ** 2))[dq])
                dn = float(np.sum(dq))
                # ordered sum counts each off-diagonal twice and each diagonal once
                P_c = 0.5 * (sumP + dP)
                Pnc_c = 0.5 * (sumPnc + dPnc)
                U_c = 0.5 * (sumU - dUo) + dU
                B_c = 0.5 * (sumB + dB)
                nq_c = 0.5 * (nq + dn)
            else:
                P_c, Pnc_c, U_c, B_c, nq_c = sumP, sumPnc, sumU, sumB, nq
            flop_eval = P_c * (40 + 12 * nr + nr * (18 * V + 3 * nf)) + 2 * nf * Pnc_c
            key = (lb_[0], lb_[1], lk_[0], lk_[1])
            out[key] = dict(quartets=nq_c, prim_quartets=P_c, flop_eval=flop_eval, flop_dig=12 * U_c,
                            bytes_dig=B_c, unique_ao_quartets=U_c)
    tot = {k: sum(v[k] for v in out.values()) for k in
           ("quartets", "prim_quartets", "flop_eval", "flop_dig", "bytes_dig", "unique_ao_quartets")}
    return {"classes": out, "total": tot}
