"""The GPU adjoint (stfem_adjoint_gradient through the C ABI) against the
FP64 CPU adjoint of the oracle and against central finite differences of the
GPU misfit, and wrapped in a torch.autograd.Function.

Problem: heat slabs of TimeIntegratorHeat::solve (time_integrators.h:157-180
of nlsmrg/dealii-stfem) with the manufactured source, zero initial values,
observations = forward solution at a seeded log-uniform coefficient in
[0.1, 10] plus 1% Gaussian noise, parameters = per-cell log coefficients.
The full-size case is 24^3 cells (3^3 coarse cells refined three times), Q2,
dG(1), 8 slabs, 16 seeded directions with h = 1e-6.
"""
import numpy as np
import pytest

from helpers import blocks_to_nodes, log_uniform

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
import stfem_amd as st  # noqa: E402

DEV = "cuda:0"


def to_nodes(traj_blocks, n_slabs, nt):
    m = traj_blocks.size // n_slabs
    return np.concatenate([blocks_to_nodes(traj_blocks[k * m:(k + 1) * m], nt)
                           for k in range(n_slabs)])


def noisy(traj, seed):
    rng = np.random.default_rng(seed)
    sigma = 0.01 * np.sqrt(np.mean(traj ** 2))
    return traj + sigma * rng.standard_normal(traj.size)


@pytest.mark.parametrize("spec,n_slabs", [
    (dict(dim=2, refine=3, p=2, type=2, r=1, nts=1, r_min=0, tau=0.05), 4),
    (dict(dim=2, refine=2, p=3, type=2, r=2, nts=2, r_min=0, tau=0.1), 3),
    (dict(dim=3, subdiv="3,3,3", refine=1, p=2, type=2, r=1, nts=1, r_min=0, tau=0.125), 3),
])
def test_adjoint_matches_oracle(spec, n_slabs):
    spec = dict(spec, reltol=1e-13, abstol=1e-300, maxiter=400, mg_float=0)
    o_true = orc.Problem(spec, kappa=None)
    n_cells = o_true.level_info(o_true.n_levels - 1)["n_cells"]
    kap_true = log_uniform(n_cells, 0.1, 10.0, seed=31)
    traj = orc.forward_trajectory(orc.Problem(spec, kappa=kap_true), n_slabs)
    obs = noisy(traj, 32)
    kap = np.exp(np.random.default_rng(33).uniform(-0.3, 0.3, n_cells))
    g_o, J_o = orc.adjoint_gradient(orc.Problem(spec, kappa=kap),
                                    orc.Problem(dict(spec, transpose_time=1), kappa=kap),
                                    obs, n_slabs)
    P = st.Problem(spec, kappa=kap)
    obs_g = torch.tensor(to_nodes(obs, n_slabs, P.nt), dtype=torch.float64, device=DEV)
    grad = torch.zeros(n_cells, dtype=torch.float64, device=DEV)
    print(f"[adjoint vs oracle] {spec}", flush=True)
    J, its = P.adjoint_gradient(obs_g, n_slabs, grad)
    g = grad.cpu().numpy()
    dg = np.abs(g - g_o).max() / np.abs(g_o).max()
    print(f"  misfit rel diff {abs(J - J_o) / J_o:.2e}, gradient rel diff {dg:.2e}", flush=True)
    assert abs(J - J_o) < 1e-12 * J_o, (J, J_o)
    assert dg < 1e-12, dg
    # forward trajectories agree too
    tr = torch.zeros(n_slabs * P.n_nodes * P.nt, dtype=torch.float64, device=DEV)
    P.set_cell_coefficient(kap_true)
    P.forward_trajectory(n_slabs, tr)
    ref = to_nodes(traj, n_slabs, P.nt)
    assert np.abs(tr.cpu().numpy() - ref).max() < 1e-10 * np.abs(ref).max()
    P.close()


class AdjointMisfit(torch.autograd.Function):
    """J(theta) with theta the per-cell log coefficients; the backward pass
    returns the adjoint gradient computed in the forward pass."""

    @staticmethod
    def forward(ctx, theta, problem, obs, n_slabs):
        problem.set_cell_coefficient(theta.detach().exp().cpu().numpy())
        grad = torch.zeros_like(theta)
        J, _ = problem.adjoint_gradient(obs, n_slabs, grad)
        ctx.save_for_backward(grad)
        return theta.new_tensor(J)

    @staticmethod
    def backward(ctx, gJ):
        (grad,) = ctx.saved_tensors
        return gJ * grad, None, None, None


def test_adjoint_config4_central_differences():
    """24^3 cells, Q2, dG(1), 8 slabs: gradient at log kappa = 0 through the
    autograd wrapper against central differences (h = 1e-6) along 16 seeded
    directions."""
    n_slabs = 8
    spec = dict(dim=3, subdiv="3,3,3", refine=3, p=2, type=2, r=1, nts=1, r_min=0, tau=0.125,
                reltol=1e-12, abstol=1e-300, maxiter=400, mg_float=0, vanka_factored=1)
    n_cells = 24 ** 3
    kap_true = log_uniform(n_cells, 0.1, 10.0, seed=41)
    print("[config4] observations", flush=True)
    Pt = st.Problem(spec, kappa=kap_true)
    traj = torch.zeros(n_slabs * Pt.n_nodes * Pt.nt, dtype=torch.float64, device=DEV)
    Pt.forward_trajectory(n_slabs, traj)
    Pt.close()
    rng = np.random.default_rng(42)
    sigma = 0.01 * torch.sqrt(torch.mean(traj ** 2)).item()
    obs = traj + sigma * torch.tensor(rng.standard_normal(traj.numel()), device=DEV)
    P = st.Problem(spec, kappa=np.ones(n_cells))
    theta = torch.zeros(n_cells, dtype=torch.float64, device=DEV, requires_grad=True)
    print("[config4] adjoint gradient", flush=True)
    J = AdjointMisfit.apply(theta, P, obs, n_slabs)
    J.backward()
    g = theta.grad.cpu().numpy()
    h = 1e-6
    drng = np.random.default_rng(43)
    worst = 0.0
    for k in range(16):
        d = drng.standard_normal(n_cells)
        P.set_cell_coefficient(np.exp(h * d))
        Jp, _ = P.misfit(obs, n_slabs)
        P.set_cell_coefficient(np.exp(-h * d))
        Jm, _ = P.misfit(obs, n_slabs)
        fd = (Jp - Jm) / (2 * h)
        rel = abs(fd - g @ d) / abs(g @ d)
        print(f"[config4] direction {k}: fd {fd:.10e} adjoint {g @ d:.10e} rel {rel:.2e}",
              flush=True)
        worst = max(worst, rel)
    assert worst < 1e-6, worst
    P.close()


@pytest.mark.parametrize("spec,msg", [
    (dict(dim=2, refine=2, p=2, type=1, r=2, nts=1, r_min=1, tau=0.1), "dG time"),
    (dict(dim=2, refine=2, p=2, type=1, r=2, nts=1, r_min=1, tau=0.1, problem=1), "heat problem"),
])
def test_adjoint_rejects_unsupported_discretisations(spec, msg):
    """The adjoint covers dG heat slabs; a cG or wave problem gets an error
    status (no device work), with a message naming the restriction."""
    P = st.Problem(spec)
    obs = torch.zeros(2 * P.n_nodes * P.nt, dtype=torch.float64, device=DEV)
    grad = torch.zeros(P.n_cells, dtype=torch.float64, device=DEV)
    with pytest.raises(st.StfemError, match=msg):
        P.adjoint_gradient(obs, 2, grad)
    P.close()
