"""The oracle's discrete adjoint (FP64, CPU) pinned by central finite
differences of its own forward misfit.

The slab sequence is TimeIntegratorHeat::solve (time_integrators.h:157-180
of nlsmrg/dealii-stfem) from zero initial values; the misfit is
J = 1/2 sum_n (x_n - d_n)^T (Alpha (x) M) (x_n - d_n) and the parameters are
the per-cell log coefficients.  A gradient that matches central differences
along random directions to ~1e-7 pins the adjoint equations, the backward
recursion, the cell gradient and the misfit weights together.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import log_uniform

SPEC = dict(dim=2, refine=2, p=2, type=2, r=1, nts=1, r_min=0, tau=0.1, reltol=1e-13,
            abstol=1e-300, maxiter=400, mg_float=0)
N_SLABS = 3


def observations(spec, n_cells, seed):
    kap_true = log_uniform(n_cells, 0.1, 10.0, seed=seed)
    traj = orc.forward_trajectory(orc.Problem(spec, kappa=kap_true), N_SLABS)
    rng = np.random.default_rng(seed + 1)
    sigma = 0.01 * np.sqrt(np.mean(traj ** 2))
    return traj + sigma * rng.standard_normal(traj.size)


@pytest.mark.parametrize("spec", [SPEC, dict(SPEC, nts=2, r=2, p=1)])
def test_oracle_adjoint_matches_central_differences(spec):
    n_cells = 16
    obs = observations(spec, n_cells, seed=21)
    theta = np.random.default_rng(3).uniform(-0.5, 0.5, n_cells)
    kap = np.exp(theta)
    P = orc.Problem(spec, kappa=kap)
    PT = orc.Problem(dict(spec, transpose_time=1), kappa=kap)
    g, J = orc.adjoint_gradient(P, PT, obs, N_SLABS)
    assert abs(J - orc.misfit(P, obs, N_SLABS)) < 1e-13 * J
    h = 1e-5
    rng = np.random.default_rng(8)
    for _ in range(3):
        d = rng.standard_normal(n_cells)
        Jp = orc.misfit(orc.Problem(spec, kappa=np.exp(theta + h * d)), obs, N_SLABS)
        Jm = orc.misfit(orc.Problem(spec, kappa=np.exp(theta - h * d)), obs, N_SLABS)
        fd = (Jp - Jm) / (2 * h)
        assert abs(fd - g @ d) < 1e-7 * abs(g @ d) + 1e-12 * J, (fd, g @ d)


def test_oracle_adjoint_zero_at_exact_data():
    """Noise-free data from the same coefficient: J = 0 and the gradient
    vanishes (the adjoint right-hand sides are zero)."""
    n_cells = 16
    kap = log_uniform(n_cells, 0.1, 10.0, seed=5)
    P = orc.Problem(SPEC, kappa=kap)
    PT = orc.Problem(dict(SPEC, transpose_time=1), kappa=kap)
    obs = orc.forward_trajectory(P, N_SLABS)
    g, J = orc.adjoint_gradient(P, PT, obs, N_SLABS)
    assert J < 1e-28 and np.abs(g).max() < 1e-14
