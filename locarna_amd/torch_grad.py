"""torch.autograd binding of the LocARNA-P gradient (thin layer over the C
ABI; the inside and adjoint passes run in the library's kernels).

`LogZ.apply(theta, batch)` returns log Z of every pair of `batch` for the
score parameters `theta` (NTHETA float64 offsets from the reference scores:
sigma 4x4, RIBOSUM arc-match term 16x16, indel, indel opening, struct
weight); its backward multiplies the incoming gradient with the per-pair
d log Z / d theta the adjoint pass computed.

`fit_expected_statistics` fits theta by gradient steps to the alignment
statistics of a generating theta: the negative log-likelihood of alignments
drawn from the generating parameters, in the limit of many alignments per
pair, is  sum_k log Z_k(theta) - theta . E_gen[f_k] / T  (scores are affine in
theta with feature vector f), whose gradient is
E_theta[f]/T - E_gen[f]/T = dlogZ/dtheta(theta) - dlogZ/dtheta(theta_gen).
"""
import math

import torch

from .api import NTHETA


class LogZ(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, batch):
        th = theta.detach().to("cpu", torch.float64)
        batch.pf_run(th.tolist())
        n = len(batch.pairs)
        logz = torch.tensor([math.log(batch.pf_result(k)[0]) for k in range(n)], dtype=torch.float64)
        grads = torch.tensor([batch.pf_grad(k) for k in range(n)], dtype=torch.float64)
        ctx.save_for_backward(grads)
        ctx.theta_device, ctx.theta_dtype = theta.device, theta.dtype
        return logz.to(theta.device, theta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (grads,) = ctx.saved_tensors
        g = grad_out.detach().to("cpu", torch.float64) @ grads
        return g.to(ctx.theta_device, ctx.theta_dtype), None


def generating_theta(seed=6, scale=30.0):
    """Seeded generating parameters (offsets from the reference scores)."""
    gen = torch.Generator().manual_seed(seed)
    return torch.randn(NTHETA, generator=gen, dtype=torch.float64) * scale


def expected_statistics(batch, theta):
    """Per pair d log Z / d theta at theta (= E_theta[f] / T)."""
    batch.pf_run(theta.detach().to("cpu", torch.float64).tolist())
    return torch.tensor([batch.pf_grad(k) for k in range(len(batch.pairs))], dtype=torch.float64)


def fit_expected_statistics(batch, theta_gen, steps=20, lr=2000.0, theta0=None):
    """Gradient descent on  sum_k log Z_k(theta) - theta . target_k  with
    target_k = dlogZ_k/dtheta(theta_gen).  Returns the per-step losses and
    distances |theta - theta_gen| (theta, the last iterate, as well)."""
    target = expected_statistics(batch, theta_gen).sum(0)
    theta = (theta0.clone() if theta0 is not None else torch.zeros(NTHETA, dtype=torch.float64))
    theta.requires_grad_(True)
    n = len(batch.pairs)
    losses, dist = [], []
    for _ in range(steps):
        logz = LogZ.apply(theta, batch)
        loss = logz.sum() - theta @ target
        loss.backward()
        with torch.no_grad():
            theta -= lr / n * theta.grad
            theta.grad = None
        losses.append(loss.item())
        dist.append((theta.detach() - theta_gen).norm().item())
    return losses, dist, theta.detach()
