"""torch.autograd wrapper for the in-register basket price.

`basket_call(s0, mu, sigma, ...)` returns the discounted basket-call price as
a tensor that depends differentiably on the per-asset s0, mu and sigma
tensors.  The forward pass runs the pricing kernel; the backward pass runs
the hand-written adjoint kernel (stocha_basket with greeks on), which replays
every path's block-split stream to recompute its terminal values instead of
storing paths, and scales the pathwise gradients by the incoming gradient.
Both passes use the same streams, so the gradient is exactly the derivative
of the forward estimator (common random numbers).
"""
from __future__ import annotations

import numpy as np
import torch

from . import basket_price


class BasketCall(torch.autograd.Function):
    @staticmethod
    def forward(ctx, s0, mu, sigma, corr, weights, strike, r, t, steps, n_paths, seed,
                precision):
        args = [x.detach().to("cpu", torch.float64).numpy() for x in (s0, mu, sigma)]
        corr_np = np.asarray(corr.detach().cpu().numpy() if torch.is_tensor(corr) else corr,
                             dtype=np.float64)
        w_np = np.asarray(weights.detach().cpu().numpy() if torch.is_tensor(weights) else weights,
                          dtype=np.float64)
        res = basket_price(*args, corr_np, w_np, strike, r, t, steps, n_paths, seed=seed,
                           precision=precision, greeks=False)
        ctx.save_for_backward(s0, mu, sigma)
        ctx.consts = (corr_np, w_np, strike, r, t, steps, n_paths, seed, precision)
        return torch.tensor(res["price"], dtype=s0.dtype, device=s0.device)

    @staticmethod
    def backward(ctx, grad):
        s0, mu, sigma = ctx.saved_tensors
        corr_np, w_np, strike, r, t, steps, n_paths, seed, precision = ctx.consts
        args = [x.detach().to("cpu", torch.float64).numpy() for x in (s0, mu, sigma)]
        res = basket_price(*args, corr_np, w_np, strike, r, t, steps, n_paths, seed=seed,
                           precision=precision, greeks=True)
        g = grad.detach().to("cpu", torch.float64)

        def out(v, like):
            return (g * torch.from_numpy(np.ascontiguousarray(v))).to(like.dtype).to(like.device)

        return (out(res["delta"], s0), out(res["dmu"], mu), out(res["vega"], sigma),
                None, None, None, None, None, None, None, None, None)


def basket_call(s0, mu, sigma, corr, weights, strike, r, t, steps, n_paths, seed=42,
                precision=32):
    """Differentiable basket-call price (see module docstring)."""
    return BasketCall.apply(s0, mu, sigma, corr, weights, strike, r, t, steps, n_paths, seed,
                            precision)
