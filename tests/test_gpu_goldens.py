"""The HIP path against the reference's own golden transcripts
(tests/transfer_01.output and tests/tp_01.output of nlsmrg/dealii-stfem, as
JSON fixtures) and the device error calculator against the oracle's.

The whole heat driver runs on the GPU (stfem_run_heat: slab solves with
FGMRES + space-time multigrid, extrapolated starts, ErrorCalculator), and its
DoF counts, level counts, slab counts, space-time error norms and average
iterations per slab are checked with the tolerances the oracle's own golden
tests use.
"""
import numpy as np
import pytest

from helpers import blocks_to_nodes, nodes_to_blocks, rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
import stfem_amd as st  # noqa: E402
from test_oracle_goldens import load, tp01_heat_runs, tp01_wave_runs, transfer01_runs  # noqa: E402

DEV = "cuda:0"


@pytest.mark.parametrize("spec", [dict(dim=2, refine=2, type=2, r=1, nts=2),
                                  dict(dim=2, refine=2, type=1, r=2, nts=2),
                                  dict(dim=3, refine=1, type=2, r=2, nts=1, p=3)])
def test_slab_error_matches_oracle(spec):
    o, g = orc.Problem(spec), st.Problem(spec)
    n, nt = g.n_nodes, g.nt
    rng = np.random.default_rng(29)
    x = rng.standard_normal(n * nt)
    prev = rng.standard_normal(n)
    ref = orc.slab_error(o, x, prev, 0.3)
    xg = torch.tensor(blocks_to_nodes(x, nt), dtype=torch.float64, device=DEV)
    pg = torch.tensor(prev, dtype=torch.float64, device=DEV)
    print(f"[slab error {spec}]", flush=True)
    got = g.slab_error(xg, pg, 0.3)
    assert np.allclose(got, ref, rtol=1e-12, atol=0), (got, ref)
    g.close()


def test_transfer01_transcript_on_gpu():
    gold = load("transfer_01")
    table_rows = [row for t in gold["tables"] for row in t]
    exact = 0
    for k, (spec, g, row) in enumerate(zip(transfer01_runs(), gold["runs"], table_rows)):
        print(f"[transfer_01 run {k}] {spec}", flush=True)
        P = st.Problem(spec)
        res = P.run_heat()
        assert res["rc"] == 0, (k, res)
        assert res["ndofs"] == g["dofs"] == row["s_dofs"], k
        assert P.n_levels - 1 == g["max_level"], k
        assert res["slabs"] == g["slabs"], k
        assert res["nblocks"] == row["t_dofs"], k
        for key in ("L2", "Linf", "H1"):
            assert abs(res[key] - row[key]) <= 6e-6 * abs(row[key]), (k, key, res[key], row[key])
        avg = res["total_its"] / res["slabs"]
        assert abs(avg - g["avg_its"]) <= 0.25, (k, avg, g["avg_its"])
        exact += abs(avg - g["avg_its"]) < 1e-12
        P.close()
    assert exact >= 30, exact


def test_tp01_transcript_on_gpu():
    exact, n = 0, 0
    for spec, g, row in tp01_heat_runs(3):
        print(f"[tp_01] {spec}", flush=True)
        P = st.Problem(spec)
        res = P.run_heat()
        assert res["rc"] == 0, (spec, res)
        assert res["ndofs"] == g["dofs"] == row["s_dofs"]
        assert P.n_levels - 1 == g["max_level"]
        assert res["slabs"] == g["slabs"]
        assert res["nblocks"] == row["t_dofs"]
        for key in ("L2", "Linf", "H1"):
            assert abs(res[key] - row[key]) <= 2e-5 * abs(row[key]), (spec, key, res[key], row[key])
        avg = res["total_its"] / res["slabs"]
        assert abs(avg - g["avg_its"]) <= 0.25, (spec, avg, g["avg_its"])
        exact += abs(avg - g["avg_its"]) < 1e-12
        n += 1
        P.close()
    assert exact >= int(0.7 * n), (exact, n)


@pytest.mark.parametrize("spec", [dict(dim=2, refine=3, type=2, r=1, nts=1),
                                  dict(dim=2, refine=3, type=2, r=2, nts=4),
                                  dict(dim=2, refine=3, type=1, r=2, nts=1, r_min=1),
                                  dict(dim=2, refine=3, type=1, r=3, nts=4, r_min=1),
                                  dict(dim=3, refine=2, type=2, r=1, nts=2)])
def test_wave_slabs_match_oracle(spec):
    """TimeIntegratorWave::solve on the GPU against the oracle's, slab by
    slab: displacements and velocities of every block to 1e-12 relative and
    the FGMRES iteration counts (the right-hand side from prev_v enters from
    the second slab on, where the velocities are non-zero)."""
    spec = dict(spec, problem=1, mg_float=1)
    o, g = orc.Problem(spec), st.Problem(spec)
    n, nt = g.n_nodes, g.nt
    xo, vo = np.zeros(n * nt), np.zeros(n * nt)
    xg = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    vg = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    pu = torch.zeros(n, dtype=torch.float64, device=DEV)
    pv = torch.zeros(n, dtype=torch.float64, device=DEV)
    time = 0.0
    for k in range(3):
        pu_o, pv_o = xo[(nt - 1) * n:].copy(), vo[(nt - 1) * n:].copy()
        pu.copy_(torch.tensor(pu_o, device=DEV))
        pv.copy_(torch.tensor(pv_o, device=DEV))
        rc_o, xo, vo, its_o = orc.solve_slab_wave(o, xo, pu_o, pv_o, time)
        print(f"[wave slab {k}] {spec}", flush=True)
        rc_g, its_g = g.solve_slab_wave(xg, vg, pu, pv, time)
        assert rc_o == 0 and rc_g == 0
        assert abs(its_g - its_o) <= 1, (k, its_g, its_o)
        eu = rel_err(nodes_to_blocks(xg.cpu().numpy(), nt), xo)
        ev = rel_err(nodes_to_blocks(vg.cpu().numpy(), nt), vo)
        print(f"  its {its_g} (oracle {its_o}), rel diff u {eu:.2e} v {ev:.2e}", flush=True)
        assert eu < 1e-12 and ev < 1e-12, (k, eu, ev)
        time += spec["nts"] * g.tau
    g.close()


def test_tp01_wave_transcript_on_gpu():
    """The wave runs tf05-tf08 of tests/tp_01.output (2D, refinements 2-3)
    through the GPU driver: counts, error norms to the printed digits,
    average iterations within one per slab, most of them exact."""
    exact, n = 0, 0
    for spec, g, row in tp01_wave_runs(3):
        print(f"[tp_01 wave] {spec}", flush=True)
        P = st.Problem(spec)
        res = P.run_heat()
        assert res["rc"] == 0, (spec, res)
        assert res["ndofs"] == g["dofs"] == row["s_dofs"]
        assert P.n_levels - 1 == g["max_level"]
        assert res["slabs"] == g["slabs"]
        assert res["nblocks"] == row["t_dofs"]
        for key in ("L2", "Linf", "H1"):
            assert abs(res[key] - row[key]) <= 2e-5 * abs(row[key]), (spec, key, res[key], row[key])
        avg = res["total_its"] / res["slabs"]
        assert abs(avg - g["avg_its"]) <= 1.0, (spec, avg, g["avg_its"])
        exact += abs(avg - g["avg_its"]) < 1e-12
        n += 1
        P.close()
    assert exact >= int(0.7 * n), (exact, n)


def test_wave_layered_pulse_matches_oracle():
    """The 3D wave workload's setting at 4^3 cells (Q2, cG(2), c^2 = 1 | 4
    split at z = 0.5, Gaussian pulse initial displacement, zero velocity, no
    source): three slabs against the oracle's TimeIntegratorWave::solve,
    displacements and velocities of every block to 1e-12 relative."""
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
    import bench
    spec, X, kappa, centre = bench.workload_wave(2)
    spec = dict(spec, mg_float=1)
    o = orc.Problem(dict(spec, rhs=1), coords=X, kappa=kappa)
    g = st.Problem(spec, vertices=X, kappa=kappa)
    n, nt = g.n_nodes, g.nt
    u0 = bench.wave_initial(g, centre)
    assert np.abs(u0).max() > 0.1
    xo, vo = np.zeros(n * nt), np.zeros(n * nt)
    xg = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    vg = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    pu_o, pv_o = u0.copy(), np.zeros(n)
    for k in range(3):
        pu = torch.tensor(pu_o, dtype=torch.float64, device=DEV)
        pv = torch.tensor(pv_o, dtype=torch.float64, device=DEV)
        rc_o, xo, vo, its_o = orc.solve_slab_wave(o, xo, pu_o, pv_o, k * g.tau)
        print(f"[wave pulse slab {k}]", flush=True)
        rc_g, its_g = g.solve_slab_wave(xg, vg, pu, pv, k * g.tau)
        assert rc_o == 0 and rc_g == 0
        assert abs(its_g - its_o) <= 1, (k, its_g, its_o)
        eu = rel_err(nodes_to_blocks(xg.cpu().numpy(), nt), xo)
        ev = rel_err(nodes_to_blocks(vg.cpu().numpy(), nt), vo)
        print(f"  its {its_g} (oracle {its_o}), rel diff u {eu:.2e} v {ev:.2e}", flush=True)
        assert eu < 1e-12 and ev < 1e-12, (k, eu, ev)
        pu_o, pv_o = xo[(nt - 1) * n:].copy(), vo[(nt - 1) * n:].copy()
    g.close()
