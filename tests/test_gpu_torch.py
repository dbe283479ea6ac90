"""The torch.autograd binding of d log Z / d theta: backward against central
finite differences of the GPU's own log Z along random directions, and the
fitting loops (expected statistics of a generating theta, and alignments
sampled at it) decreasing their loss and their distance to the generating
theta."""
import pytest
import torch

from locarna_amd import api, synth
from locarna_amd.torch_grad import (LogZ, fit_expected_statistics, fit_sampled_alignments, generating_theta,
                                    sampled_statistics)

pytestmark = pytest.mark.gpu


def _batch(npairs):
    pps, pairs, span = synth.workload("pf16x150", limit_pairs=npairs)
    pa = api.locarna_p_params(max_bp_span=span)
    P = [api.Pair(pps[a], pps[b], pa) for a, b in pairs]
    return api.Batch(P)


def test_backward_matches_finite_differences():
    B = _batch(2)
    gen = torch.Generator().manual_seed(11)
    theta = (torch.randn(api.NTHETA, generator=gen, dtype=torch.float64) * 5).requires_grad_(True)
    w = torch.tensor([0.7, -1.3], dtype=torch.float64)
    (LogZ.apply(theta, B) * w).sum().backward()
    g = theta.grad.clone()
    h = 1e-3
    for trial in range(3):
        v = torch.randn(api.NTHETA, generator=gen, dtype=torch.float64)
        v /= v.norm()
        with torch.no_grad():
            up = (LogZ.apply(theta + h * v, B) * w).sum().item()
            dn = (LogZ.apply(theta - h * v, B) * w).sum().item()
        fd = (up - dn) / (2 * h)
        assert abs(fd - (g @ v).item()) <= 1e-6 * max(1.0, abs(fd)), (trial, fd, (g @ v).item())
    B.destroy()


def test_fitting_moves_towards_generating_theta():
    """Newton steps with the sampled diagonal Hessian recover at least half
    of the identifiable part of a generating theta from its expected
    statistics; plain gradient descent with the same number of steps moves
    far less."""
    B = _batch(4)
    theta_gen = generating_theta(seed=6, scale=30.0)
    losses, dist, _ = fit_expected_statistics(B, theta_gen, steps=8)
    print("newton losses", losses, "dist", dist, flush=True)
    assert losses[-1] < losses[0]
    assert dist[-1] <= 0.5 * dist[0]
    gl, gd, _ = fit_expected_statistics(B, theta_gen, steps=8, lr=2000.0, method="gd")
    print("gd losses", gl, "dist", gd, flush=True)
    assert dist[-1] < gd[-1]
    B.destroy()


def test_fit_to_sampled_alignments():
    B = _batch(4)
    theta_gen = generating_theta(seed=6, scale=30.0)
    # seeded: the same seed gives the same target, another seed another one
    t1 = sampled_statistics(B, theta_gen, samples=32, seed=3)
    t2 = sampled_statistics(B, theta_gen, samples=32, seed=3)
    t3 = sampled_statistics(B, theta_gen, samples=32, seed=4)
    assert torch.equal(t1, t2) and not torch.equal(t1, t3)
    losses, dist, _ = fit_sampled_alignments(B, theta_gen, samples=256, seed=6, steps=6)
    print("losses", losses, "dist", dist, flush=True)
    assert losses[-1] < losses[0]
    assert dist[-1] < dist[0]
    B.destroy()
