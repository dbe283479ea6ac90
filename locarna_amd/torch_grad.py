"""torch.autograd binding of the LocARNA-P gradient (thin layer over the C
ABI; the inside and adjoint passes run in the library's kernels).

`LogZ.apply(theta, batch)` returns log Z of every pair of `batch` for the
score parameters `theta` (NTHETA float64 offsets from the reference scores:
sigma 4x4, RIBOSUM arc-match term 16x16, indel, indel opening, struct
weight); its backward multiplies the incoming gradient with the per-pair
d log Z / d theta the adjoint pass computed.

`fit_sampled_alignments` fits theta by gradient steps to alignments drawn
from a seeded generating theta (the library's stochastic traceback,
lcg_batch_pf_sample): scores are affine in theta with feature vector f, so
the negative log-likelihood of the sampled alignments of pair k is
log Z_k(theta) - theta . mean_k(f) / T up to a constant, with gradient
E_theta[f]/T - mean_k(f)/T = dlogZ/dtheta(theta) - mean_k(f)/T.
`fit_expected_statistics` is its limit of infinitely many samples (target
dlogZ/dtheta at the generating theta).
"""
import concurrent.futures
import math
import os

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


def sampled_statistics(batch, theta_gen, samples=64, seed=6, temperature=None):
    """Per pair mean(f) / T of `samples` alignments drawn at theta_gen."""
    batch.pf_run(theta_gen.detach().to("cpu", torch.float64).tolist())
    T = float(temperature if temperature is not None else batch.pairs[0].params.temperature_alipf)

    def one(k):
        mean, _ = batch.pf_sample(k, seed * 1000003 + k, samples)
        return [x / T for x in mean]

    # the sampler reads the fetched host results only; pairs run on threads
    with concurrent.futures.ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        rows = list(ex.map(one, range(len(batch.pairs))))
    return torch.tensor(rows, dtype=torch.float64)


def fit_sampled_alignments(batch, theta_gen, samples=64, seed=6, steps=20, lr=2000.0, theta0=None):
    """Gradient descent on  sum_k log Z_k(theta) - theta . mean_k(f) / T  for
    `samples` alignments per pair drawn at theta_gen (seeded).  Returns the
    per-step losses and distances |theta - theta_gen| and the last theta."""
    return fit_to_target(batch, theta_gen, sampled_statistics(batch, theta_gen, samples, seed).sum(0), steps, lr, theta0)


def fit_expected_statistics(batch, theta_gen, steps=20, lr=2000.0, theta0=None):
    """Gradient descent on  sum_k log Z_k(theta) - theta . target_k  with
    target_k = dlogZ_k/dtheta(theta_gen).  Returns the per-step losses and
    distances |theta - theta_gen| (theta, the last iterate, as well)."""
    return fit_to_target(batch, theta_gen, expected_statistics(batch, theta_gen).sum(0), steps, lr, theta0)


def fit_to_target(batch, theta_gen, target, steps=20, lr=2000.0, theta0=None):
    """`steps` gradient steps on  sum_k log Z_k(theta) - theta . target
    (target: summed per-pair statistics / T); returns the losses, the
    distances |theta - theta_gen| after each step and the last theta."""
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
