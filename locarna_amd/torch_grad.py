"""torch.autograd binding of the LocARNA-P gradient (thin layer over the C
ABI; the inside and adjoint passes run in the library's kernels).

`LogZ.apply(theta, batch)` returns log Z of every pair of `batch` for the
score parameters `theta` (NTHETA float64 offsets from the reference scores:
sigma 4x4, RIBOSUM arc-match term 16x16, indel, indel opening, struct
weight); its backward multiplies the incoming gradient with the per-pair
d log Z / d theta the adjoint pass computed.

`fit_sampled_alignments` fits theta to alignments drawn from a seeded
generating theta (the library's stochastic traceback, lcg_batch_pf_sample):
scores are affine in theta with feature vector f, so the negative
log-likelihood of the sampled alignments of pair k is log Z_k(theta) -
theta . mean_k(f) / T up to a constant, with gradient E_theta[f]/T -
mean_k(f)/T = dlogZ/dtheta(theta) - mean_k(f)/T and Hessian Cov_theta(f)/T^2.
The default steps are Newton steps with the diagonal of that Hessian
estimated by the sampler.  `fit_expected_statistics` is the limit of
infinitely many samples (target dlogZ/dtheta at the generating theta).
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


def fisher_diagonal(batch, samples=64, seed=0, temperature=None):
    """Diagonal of the Hessian of sum_k log Z_k at the parameters of the
    batch's last pf_run: sum_k Var_k(f) / T^2 (log Z is the cumulant
    function of the feature vector f / T), estimated from `samples`
    alignments per pair of the stochastic traceback."""
    T = float(temperature if temperature is not None else batch.pairs[0].params.temperature_alipf)

    def one(k):
        _, var = batch.pf_sample(k, seed * 1000003 + k, samples)
        return var

    with concurrent.futures.ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        rows = list(ex.map(one, range(len(batch.pairs))))
    return torch.tensor(rows, dtype=torch.float64).sum(0) / (T * T)


def fit_sampled_alignments(batch, theta_gen, samples=64, seed=6, steps=20, lr=0.7, theta0=None, method="newton",
                           prior_scale=30.0):
    """Fits  sum_k log Z_k(theta) - theta . mean_k(f) / T  for `samples`
    alignments per pair drawn at theta_gen (seeded), with the Gaussian prior
    of the generating parameters (standard deviation prior_scale, see
    generating_theta) as a ridge term |theta|^2 / (2 prior_scale^2 samples):
    the maximum a posteriori estimate, which keeps components that the
    samples barely observe near the prior instead of at the noise of their
    few observations.  Returns the per-step losses, distances |theta -
    theta_gen| over the identifiable components and the last theta (see
    fit_to_target)."""
    target = sampled_statistics(batch, theta_gen, samples, seed).sum(0)
    ridge = 1.0 / (prior_scale * prior_scale * samples) if prior_scale else 0.0
    return fit_to_target(batch, theta_gen, target, steps, lr, theta0, method=method, seed=seed, ridge=ridge)


def fit_expected_statistics(batch, theta_gen, steps=20, lr=0.7, theta0=None, method="newton"):
    """Fits  sum_k log Z_k(theta) - theta . target_k  with target_k =
    dlogZ_k/dtheta(theta_gen) (infinitely many samples)."""
    return fit_to_target(batch, theta_gen, expected_statistics(batch, theta_gen).sum(0), steps, lr, theta0,
                         method=method)


def identifiable(batch, theta_gen, samples=64, seed=1):
    """Components of theta that the data determine: those whose feature
    varies under the distribution at theta_gen (a component whose feature
    never occurs, e.g. an arc-match term of base-pair types no input arc
    has, has zero gradient and curvature, and no fit can recover it)."""
    batch.pf_run(theta_gen.detach().to("cpu", torch.float64).tolist())
    return fisher_diagonal(batch, samples, seed) > 0


def fit_to_target(batch, theta_gen, target, steps=20, lr=0.7, theta0=None, method="newton", samples=64, seed=0,
                  damping=1e-3, max_step=50.0, ridge=0.0):
    """`steps` steps on the convex loss  sum_k log Z_k(theta) - theta . target
    + ridge |theta|^2 / 2  (target: summed per-pair statistics / T).

    method "newton": theta -= lr g / (H + damping mean(H)) with g the adjoint
    gradient and H the sampled diagonal Hessian (fisher_diagonal) at the
    current theta, each step at most max_step per component; a step that
    raised the loss is halved (theta moves back half way) and lr with it (the
    diagonal Hessian ignores the coupling of the parameters, so full steps can
    overshoot); method "gd": theta -= lr / npairs g.  Returns the losses, the
    distances |theta - theta_gen| over the identifiable components before the
    first step and after each step, and the last theta."""
    theta = (theta0.clone() if theta0 is not None else torch.zeros(NTHETA, dtype=torch.float64))
    mask = identifiable(batch, theta_gen)
    theta.requires_grad_(True)
    n = len(batch.pairs)
    losses, dist = [], [(theta.detach() - theta_gen)[mask].norm().item()]
    prev_theta, prev_loss = None, None
    for step in range(steps):
        logz = LogZ.apply(theta, batch)
        loss = logz.sum() - theta @ target + 0.5 * ridge * (theta @ theta)
        if method == "newton" and prev_loss is not None and loss.item() > prev_loss:
            with torch.no_grad():
                theta.copy_(0.5 * (theta + prev_theta))
            lr *= 0.5
            losses.append(loss.item())
            dist.append((theta.detach() - theta_gen)[mask].norm().item())
            continue
        loss.backward()
        prev_theta, prev_loss = theta.detach().clone(), loss.item()
        with torch.no_grad():
            g = theta.grad
            if method == "newton":
                H = fisher_diagonal(batch, samples, seed + 7919 * (step + 1))
                scale = H[H > 0].mean() if bool((H > 0).any()) else torch.tensor(1.0, dtype=torch.float64)
                d = torch.where(H > 0, g / (H + ridge + damping * scale), torch.zeros_like(g))
                theta -= lr * d.clamp(-max_step, max_step)
            else:
                theta -= lr / n * g
            theta.grad = None
        losses.append(loss.item())
        dist.append((theta.detach() - theta_gen)[mask].norm().item())
    return losses, dist, theta.detach()
