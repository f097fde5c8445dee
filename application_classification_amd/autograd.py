"""torch.autograd binding of the matcher's adjoint (ac_set_weights / ac_run / ac_backward through ctypes).

    m = Matcher(0); m.set_option(AC_OPT_TAPE, 4); m.set_data(data); m.set_pattern(pattern)
    wv = torch.ones(dV, requires_grad=True); we = torch.ones(dE, requires_grad=True)
    mu = match_log_probs(m, wv, we)          # dn x pn log-probabilities (float32, CPU tensor)
    loss = planted_cross_entropy(mu, vmap); loss.backward()   # wv.grad, we.grad

The forward runs the fused HIP matcher with the weights in its cdist and records the tape; the backward hands
dL/dMU to the hand-written adjoint.  A backward must follow its own forward on the same Matcher (the tape holds
the last run only).
"""
from __future__ import annotations

import numpy as np
import torch

from .matcher import AC_OPT_TAPE, Matcher


class MatchFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, wv, we, matcher: Matcher, n_iter: int):
        if matcher.get_option(AC_OPT_TAPE) <= 0:
            raise RuntimeError("MatchFunction needs a Matcher with AC_OPT_TAPE > 0")
        matcher.set_weights(wv.detach().cpu().numpy(), we.detach().cpu().numpy())
        mu = matcher.run(n_iter)
        ctx.matcher = matcher
        ctx.devices = (wv.device, we.device)
        return torch.from_numpy(mu)

    @staticmethod
    def backward(ctx, grad_mu):
        gwv, gwe = ctx.matcher.backward(grad_mu.detach().cpu().numpy())
        return torch.from_numpy(gwv).to(ctx.devices[0]), torch.from_numpy(gwe).to(ctx.devices[1]), None, None


def match_log_probs(matcher: Matcher, wv: torch.Tensor, we: torch.Tensor, n_iter: int = -1) -> torch.Tensor:
    """MU (dn x pn) of the matcher with per-dimension cdist weights wv, we; differentiable in wv and we."""
    return MatchFunction.apply(wv, we, matcher, n_iter)


def planted_cross_entropy(mu: torch.Tensor, vmap) -> torch.Tensor:
    """Mean over pattern vertices p of -log P(v = vmap[p] | p) = -MU[vmap[p]][p] (MU is a log-softmax over v)."""
    idx = torch.as_tensor(np.asarray(vmap, np.int64))
    return -mu[idx, torch.arange(mu.shape[1])].mean()
