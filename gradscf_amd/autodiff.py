"""Automatic differentiation of direct J/K (torch.autograd).

Mirrors the differentiation rules of the reference's direct-J/K primitive:

  * density: J and K are linear in D and self-adjoint on nonsymmetric AO
    densities, so the JVP applies the same operator to the tangent and the
    VJP applies it to the cotangent, J to the J cotangent and K to the K
    cotangent (src/gradscf/integrals/backends/native/jk.py:284-307, the
    transpose rule `_density_transpose` at :385-391); a density cotangent is
    itself produced by the differentiable operator, so Hessian-vector
    products and mixed density/coordinate second derivatives follow;
  * coordinates: the coordinate VJP of sum(cot_J * J(D)) + sum(cot_K * K(D))
    over every atm-row coordinate (nuclei and basis centres), the reference's
    `_direct_geometry_value` rule (jk.py:321-382) on its geometry VJP target
    (csrc/direct_jk_geometry.cc:122-182); unscreened plans only, with the
    reference's message "Screened direct J/K geometry AD requires cutoff=0";
  * env: the environment vector's exponent and (normalised) coefficient
    slots (gscf_jk_param_vjp) plus its coordinate slots (the coordinate VJP
    scattered to PTR_COORD), i.e. an env-sized gradient like the reference's
    ENV-level geometry VJP.  The reference rejects exponent/coefficient
    tangents (jk.py:213-249); here they are implemented, so gradients reach
    raw exponents and coefficients through a differentiable packing
    (packing.pack_torch) and the normalisation chain.

Second derivatives with respect to coordinates or basis parameters are not
implemented: differentiating a coordinate or env gradient raises.
"""
from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np
import torch

from .packing import PTR_COORD


@dataclass(frozen=True)
class JKOperator:
    """One direct J/K operator: a cached device plan plus the operator selection."""
    plan: object            # jk.DirectJKPlan
    with_j: bool
    with_k: bool

    @property
    def nops(self) -> int:
        return int(self.with_j) + int(self.with_k)

    def select(self, with_j: bool, with_k: bool) -> "JKOperator":
        return replace(self, with_j=with_j, with_k=with_k)

    def apply(self, density: torch.Tensor) -> torch.Tensor:
        return self.plan.jk(density.detach().contiguous(), with_j=self.with_j, with_k=self.with_k)


def _split(op: JKOperator, g: torch.Tensor):
    """(cot_J, cot_K) of an output cotangent (nops, nb, n, n)."""
    gj = g[0] if op.with_j else None
    gk = g[-1] if op.with_k else None
    return gj, gk


def _coordinate_vjp(op: JKOperator, density, gj, gk) -> torch.Tensor:
    if op.plan.cutoff > 0:
        raise NotImplementedError("Screened direct J/K geometry AD requires cutoff=0")
    total = None
    for b in range(density.shape[0]):
        g = op.plan.jk_geometry_vjp(density[b], None if gj is None else gj[b], None if gk is None else gk[b])
        total = g if total is None else total + g
    return total


def _param_vjp(op: JKOperator, density, gj, gk) -> torch.Tensor:
    if op.plan.cutoff > 0:
        raise NotImplementedError("Screened direct J/K basis-parameter AD requires cutoff=0")
    total = None
    for b in range(density.shape[0]):
        g = op.plan.jk_param_vjp(density[b], None if gj is None else gj[b], None if gk is None else gk[b])
        total = g if total is None else total + g
    return total


class _FirstOrder(torch.autograd.Function):
    """Coordinate and env gradients of sum(cot_J J(D)) + sum(cot_K K(D)):
    (natm, 3) rows and an env-sized vector.  Not differentiable again."""

    @staticmethod
    def forward(ctx, op, want_rows, want_env, density, gj, gk):
        d = density.detach()
        gj = None if gj is None else gj.detach()
        gk = None if gk is None else gk.detach()
        rows = _coordinate_vjp(op, d, gj, gk)
        env = d.new_zeros(0)
        if want_env:
            slots = torch.as_tensor(op.plan.atm[:, PTR_COORD].astype(np.int64), device=d.device)
            idx = (slots[:, None] + torch.arange(3, device=d.device)[None, :]).reshape(-1)
            env = _param_vjp(op, d, gj, gk).index_add(0, idx, rows.reshape(-1))
        if not want_rows:
            rows = d.new_zeros(0)
        return rows, env

    @staticmethod
    def backward(ctx, *grads):
        raise NotImplementedError("Second derivatives of direct J/K with respect to coordinates or basis "
                                  "parameters are not implemented.")


class DirectJKFunction(torch.autograd.Function):
    """out (nops, nb, n, n) = [J(D), K(D)] of real densities D (nb, n, n) on the
    GPU, differentiable in D, in the atm-row coordinates (natm, 3) and in the
    env vector (exponent, coefficient and coordinate slots).

    coordinates / env are optional tensors that only carry gradients: the
    values the operator uses are the ones its plan was built from."""

    @staticmethod
    def forward(ctx, op: JKOperator, density, coordinates=None, env=None):
        ctx.op = op
        ctx.save_for_backward(density, coordinates, env)
        return op.apply(density)

    @staticmethod
    def backward(ctx, g):
        op = ctx.op
        density, coordinates, env = ctx.saved_tensors
        gj, gk = _split(op, g)
        grad_d = grad_c = grad_e = None
        if ctx.needs_input_grad[1]:
            # self-adjoint maps: J's cotangent goes through J, K's through K;
            # computed by this Function again, so it stays differentiable
            parts = []
            if gj is not None:
                parts.append(DirectJKFunction.apply(op.select(True, False), gj, coordinates, env)[0])
            if gk is not None:
                parts.append(DirectJKFunction.apply(op.select(False, True), gk, coordinates, env)[0])
            grad_d = parts[0] if len(parts) == 1 else parts[0] + parts[1]
        want_rows = ctx.needs_input_grad[2]
        want_env = ctx.needs_input_grad[3]
        if want_rows or want_env:
            rows, env_g = _FirstOrder.apply(op, want_rows, want_env, density, gj, gk)
            if want_rows:
                grad_c = rows.to(coordinates.dtype).to(coordinates.device)
            if want_env:
                grad_e = env_g.to(env.dtype).to(env.device)
        return None, grad_d, grad_c, grad_e

    @staticmethod
    def jvp(ctx, dop, ddensity, dcoordinates, denv):
        if dcoordinates is not None or denv is not None:
            raise NotImplementedError("Forward-mode coordinate and basis-parameter derivatives of direct J/K "
                                      "are not implemented; use reverse mode.")
        op = ctx.op
        if ddensity is None:
            return None
        return op.apply(ddensity)


def apply_jk(op: JKOperator, density: torch.Tensor, coordinates=None, env=None) -> torch.Tensor:
    """Differentiable [J, K] of a real density batch (nb, n, n)."""
    return DirectJKFunction.apply(op, density, coordinates, env)


def needs_grad(*tensors) -> bool:
    """True when any tensor takes part in reverse-mode (requires_grad under
    grad mode) or forward-mode (a dual tensor with a tangent) differentiation."""
    import torch.autograd.forward_ad as fwad
    for t in tensors:
        if not isinstance(t, torch.Tensor):
            continue
        if t.requires_grad and torch.is_grad_enabled():
            return True
        if fwad.unpack_dual(t).tangent is not None:
            return True
    return False
