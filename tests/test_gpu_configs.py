"""GPU parity for the benchmark configurations and the solver options the
parity suite's generic cases leave out, against the oracle on the same
seeded inputs, plus the failure paths of the slab solve.

* Config 1 as specified: 2D heat, 16 x 16 cells, Q2, dG(1), T = 1 in 16
  slabs, k = 1, u = sin(pi x) sin(pi y) sin(pi t) (frequency 0.5), two
  smoothing steps, FGMRES to 1e-10, all FP64: every slab against the oracle
  (identical iteration counts, solutions to 1e-12).
* The 64^3 headline's discretisation at 4^3 cells: Q4, dG(3), tau = 1/4,
  vertices perturbed by 0.1h, cellwise log-uniform kappa in [1e-3, 1e3],
  Gaussian bumps, arithmetic-mean coarse coefficients, factored Vanka, two
  smoothing steps; the plain FP64 multigrid leg (the FP64 factored sweep
  kernel) and the FP32 leg, three slabs each.
* The 32^3 workload's discretisation (Q3, dG(2), 0.2h, [1e-2, 1e2]) with
  the Gaussian-bump source, FP64 and FP32 multigrid.
* extrapolate = 0 (time_integrators.h:96-104) for heat and wave, and the
  L2 time restriction (restrictIsTransposeProlongate = false, gmg.h:101-113).
* Failure paths: the iteration limit (NoConvergence, time_integrators.h:
  172-179) gives STFEM_ERR_NO_CONVERGENCE, a NaN in the right-hand side
  STFEM_ERR_NOT_FINITE; the one-pass atomic operator equals the two-pass one.
* A reference parameter file (tests/json/tf01.json) drives the GPU heat run
  exactly as the equivalent key=value spec; the phase timers count calls.

The parity slab solves run FGMRES to 1e-12 (the library default) so that
the GPU and oracle solutions are compared beyond the solver tolerance.
"""
import json
import os

import numpy as np
import pytest

from helpers import blocks_to_nodes, box_vertices, log_uniform, nodes_to_blocks, rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
import stfem_amd as st  # noqa: E402

DEV = "cuda:0"
HERE = os.path.dirname(os.path.abspath(__file__))


def run_slabs(spec, n_slabs, X=None, kap=None, oracle_extra=None, source_key=True):
    """n_slabs heat slabs on the GPU and in the oracle from a zero state;
    the library's source key maps to the oracle's rhs key."""
    ospec = dict(spec, **(oracle_extra or {}))
    if "source" in spec:
        ospec["rhs"] = spec["source"]
    o = orc.Problem(ospec, coords=X, kappa=kap)
    g = st.Problem(spec, vertices=X, kappa=kap)
    n, nt = g.n_nodes, g.nt
    slab = spec.get("nts", 1) * g.tau
    xo = np.zeros(n * nt)
    xg = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    pg = torch.zeros(n, dtype=torch.float64, device=DEV)
    out = []
    for k in range(n_slabs):
        po = xo[(nt - 1) * n:].copy()
        pg.copy_(torch.tensor(po, device=DEV))
        rc_o, xo, its_o = orc.solve_slab(o, xo, po, k * slab)
        print(f"[slabs {spec}] slab {k}", flush=True)
        rc_g, its_g = g.solve_slab(xg, pg, k * slab)
        err = rel_err(nodes_to_blocks(xg.cpu().numpy(), nt), xo)
        print(f"  its {its_g} (oracle {its_o}), rel diff {err:.2e}", flush=True)
        out.append((rc_g, rc_o, its_g, its_o, err))
    g.close()
    return out


def check(results, fp64):
    for k, (rc_g, rc_o, its_g, its_o, err) in enumerate(results):
        assert rc_g == 0 and rc_o == 0, (k, rc_g, rc_o)
        if fp64:
            assert its_g == its_o, (k, its_g, its_o)
        else:
            assert abs(its_g - its_o) <= 1, (k, its_g, its_o)
        assert err < 1e-12, (k, err)


def test_config1_all_slabs():
    spec = dict(dim=2, refine=4, p=2, type=2, r=1, nts=1, tau=1.0 / 16, frequency=0.5,
                smoothing_steps=2, reltol=1e-10, abstol=1e-300)
    check(run_slabs(spec, 16), fp64=True)
    # the whole heat driver: errors and iterations of the oracle's run
    ro = orc.Problem(spec).run_heat()
    g = st.Problem(spec)
    rg = g.run_heat()
    g.close()
    assert rg["rc"] == 0 and rg["slabs"] == ro["slabs"] == 16
    assert rg["its"] == ro["its"]
    for key in ("L2", "Linf", "H1"):
        assert abs(rg[key] - ro[key]) <= 1e-10 * abs(ro[key]), (key, rg[key], ro[key])


def headline_small(mg_float):
    nc = [4] * 3
    X = box_vertices(3, nc, perturb=0.1, seed=2026)
    kap = log_uniform(64, 1e-3, 1e3, seed=2027)
    spec = dict(dim=3, refine=2, p=4, type=2, r=3, nts=1, tau=0.25, source=3, n_bumps=4, seed=7,
                coarse_coef=3, vanka_factored=1, smoothing_steps=2, mg_float=mg_float)
    return spec, X, kap


@pytest.mark.parametrize("mg_float", [0, 1])
def test_headline_discretisation_at_4cube(mg_float):
    spec, X, kap = headline_small(mg_float)
    extra = dict(vanka_invert_T=0) if mg_float else None
    check(run_slabs(spec, 3, X, kap, oracle_extra=extra), fp64=not mg_float)


@pytest.mark.parametrize("mg_float", [0, 1])
def test_bump_source_q3_dg2(mg_float):
    nc = [4] * 3
    X = box_vertices(3, nc, perturb=0.2, seed=2024)
    kap = log_uniform(64, 1e-2, 1e2, seed=2025)
    spec = dict(dim=3, refine=2, p=3, type=2, r=2, nts=1, r_min=0, tau=0.125, source=3,
                n_bumps=4, seed=7, vanka_factored=1, smoothing_steps=2, mg_float=mg_float)
    extra = dict(vanka_invert_T=0) if mg_float else None
    check(run_slabs(spec, 3, X, kap, oracle_extra=extra), fp64=not mg_float)


def test_heat_without_extrapolation():
    spec = dict(dim=2, refine=3, p=2, type=2, r=1, nts=2, r_min=0, extrapolate=0)
    check(run_slabs(spec, 3), fp64=True)
    # the start vector matters once the solve is cut short: two iterations
    # from a nonzero previous slab, with and without extrapolation, differ,
    # and each matches the oracle's (which also stops at the limit)
    rng = np.random.default_rng(6)
    outs = []
    for ex in (0, 1):
        sp2 = dict(spec, extrapolate=ex, maxiter=2)
        o, g = orc.Problem(sp2), st.Problem(sp2)
        n, nt = g.n_nodes, g.nt
        prev = rng.standard_normal(n) if ex == 0 else prev
        rc_o, xo, its_o = orc.solve_slab(o, np.zeros(n * nt), prev, 0.0)
        xg = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
        print(f"[extrapolate={ex}, maxiter=2]", flush=True)
        rc_g, its_g = g.solve_slab(xg, torch.tensor(prev, device=DEV), 0.0)
        assert rc_o != 0 and rc_g == -3 and its_g == its_o == 2, (rc_o, rc_g, its_g, its_o)
        assert rel_err(nodes_to_blocks(xg.cpu().numpy(), nt), xo) < 1e-12
        outs.append(xo)
        g.close()
    assert rel_err(outs[0], outs[1]) > 1e-6


def test_wave_without_extrapolation():
    spec = dict(dim=2, refine=2, p=2, type=1, r=2, nts=2, r_min=1, problem=1, extrapolate=0)
    o, g = orc.Problem(spec), st.Problem(spec)
    n, nt = g.n_nodes, g.nt
    rng = np.random.default_rng(3)
    pu0, pv0 = rng.standard_normal(n), rng.standard_normal(n)
    xo = np.zeros(n * nt)
    rc_o, xo, vo, its_o = orc.solve_slab_wave(o, xo, pu0, pv0, 0.0)
    xg = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    vg = torch.zeros_like(xg)
    print("[wave extrapolate=0]", flush=True)
    rc_g, its_g = g.solve_slab_wave(xg, vg, torch.tensor(pu0, device=DEV),
                                    torch.tensor(pv0, device=DEV), 0.0)
    assert rc_o == 0 and rc_g == 0 and its_g == its_o, (its_g, its_o)
    assert rel_err(nodes_to_blocks(xg.cpu().numpy(), nt), xo) < 1e-12
    assert rel_err(nodes_to_blocks(vg.cpu().numpy(), nt), vo) < 1e-12
    g.close()


@pytest.mark.parametrize("spec", [dict(dim=2, refine=2, p=2, type=2, r=2, nts=2, r_min=0),
                                  dict(dim=2, refine=2, p=3, type=1, r=3, nts=2, r_min=1)])
def test_l2_time_restriction(spec):
    """restrict_is_transpose=0: time restriction by the L2 restriction and
    projection matrices (fe_time.h:567-716) on every time level."""
    spec = dict(spec, restrict_is_transpose=0)
    o = orc.Problem(dict(spec, r_is_pT=0))
    g = st.Problem(spec)
    rng = np.random.default_rng(8)
    for l in range(1, g.n_levels):
        fi, ci = g.level_info(l), g.level_info(l - 1)
        if fi["mg_type"] == 3:
            continue
        xf = rng.standard_normal(fi["n_nodes"] * fi["nt"])
        yc0 = rng.standard_normal(ci["n_nodes"] * ci["nt"])
        dst = torch.tensor(blocks_to_nodes(yc0, ci["nt"]), dtype=torch.float64, device=DEV)
        src = torch.tensor(blocks_to_nodes(xf, fi["nt"]), dtype=torch.float64, device=DEV)
        print(f"[L2 restriction] level {l}", flush=True)
        g.restrict_and_add(l, dst, src)
        ref = o.level_apply(l, 3, xf, dst=yc0.copy())
        assert rel_err(nodes_to_blocks(dst.cpu().numpy(), ci["nt"]), ref) < 1e-12, l
    g.close()
    check(run_slabs(spec, 2, oracle_extra=dict(r_is_pT=0)), fp64=True)


def test_iteration_limit_reports_no_convergence():
    spec = dict(dim=2, refine=3, p=2, type=2, r=1, nts=1, maxiter=1)
    g = st.Problem(spec)
    n, nt = g.n_nodes, g.nt
    x = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    prev = torch.zeros(n, dtype=torch.float64, device=DEV)
    print("[maxiter=1]", flush=True)
    rc, its = g.solve_slab(x, prev, 0.0)
    assert rc == -3 and its == 1, (rc, its)
    assert torch.isfinite(x).all()
    g.close()


def test_non_finite_rhs_reports_not_finite():
    g = st.Problem(dict(dim=2, refine=3, p=2, type=2, r=1, nts=1))
    n, nt = g.n_nodes, g.nt
    b = torch.ones(n * nt, dtype=torch.float64, device=DEV)
    b[nt * (n // 2)] = float("nan")
    x = torch.zeros_like(b)
    print("[nan rhs]", flush=True)
    rc, its, _ = g.fgmres(x, b)
    assert rc == -5, rc
    g.close()


@pytest.mark.parametrize("mg_float", [0, 1])
def test_atomic_scatter_equals_two_pass(mg_float):
    nc = [4] * 3
    X = box_vertices(3, nc, perturb=0.2, seed=7)
    kap = log_uniform(64, 1e-2, 1e2, seed=11)
    spec = dict(dim=3, refine=2, p=3, type=2, r=2, nts=1, r_min=0, vanka_factored=1,
                mg_float=mg_float)
    a = st.Problem(spec, vertices=X, kappa=kap)
    b = st.Problem(dict(spec, atomic_scatter=1), vertices=X, kappa=kap)
    rng = np.random.default_rng(12)
    n, nt = a.n_nodes, a.nt
    x = torch.tensor(rng.standard_normal(n * nt), dtype=torch.float64, device=DEV)
    ya, yb = torch.zeros_like(x), torch.zeros_like(x)
    print(f"[atomic vs two-pass mg_float={mg_float}] fine vmult", flush=True)
    a.vmult(st.FINE, ya, x)
    b.vmult(st.FINE, yb, x)
    assert rel_err(yb.cpu().numpy(), ya.cpu().numpy()) < 1e-12
    dt = torch.float32 if mg_float else torch.float64
    for l in range(a.n_levels):
        info = a.level_info(l)
        xl = torch.tensor(rng.standard_normal(info["n_nodes"] * info["nt"]), dtype=dt, device=DEV)
        ya, yb = torch.zeros_like(xl), torch.zeros_like(xl)
        a.vmult(l, ya, xl)
        b.vmult(l, yb, xl)
        tol = 2e-6 if mg_float else 1e-12
        assert rel_err(yb.cpu().double().numpy(), ya.cpu().double().numpy()) < tol, l
    xa, xb = torch.zeros_like(x), torch.zeros_like(x)
    prev = torch.zeros(n, dtype=torch.float64, device=DEV)
    ra, ia = a.solve_slab(xa, prev, 0.0)
    rb, ib = b.solve_slab(xb, prev, 0.0)
    assert ra == 0 and rb == 0 and abs(ia - ib) <= (1 if mg_float else 0)
    assert rel_err(xb.cpu().numpy(), xa.cpu().numpy()) < 1e-12
    a.close()
    b.close()


def test_reference_parameter_file_drives_the_heat_run():
    text = open(os.path.join(HERE, "golden", "json", "tf01.json")).read()
    ref = json.loads(text)
    a = st.Problem(text)
    b_spec = dict(type=ref["timeType"], r=int(ref["feDegree"]),
                  nts=int(ref["nTimestepsAtOnce"]), refine=int(ref["refinement"]),
                  problem=ref["problemType"], dim=2, frequency=ref["frequency"])
    b = st.Problem(b_spec)
    print("[tf01.json heat run]", flush=True)
    assert st.spec_normalize(text) == st.spec_normalize(b_spec)
    ra, rb = a.run_heat(), b.run_heat()
    assert ra["rc"] == 0 and ra["its"] == rb["its"] and ra["slabs"] == rb["slabs"]
    for key in ("L2", "Linf", "H1"):
        assert abs(ra[key] - rb[key]) <= 1e-12 * abs(rb[key]), key
    a.close()
    b.close()


def test_phase_timers():
    spec = dict(dim=2, refine=3, p=2, type=2, r=1, nts=2, r_min=0, print_timing=1,
                smoothing_steps=2)
    g = st.Problem(spec)
    n, nt = g.n_nodes, g.nt
    x = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    prev = torch.ones(n, dtype=torch.float64, device=DEV)
    print("[phase timers]", flush=True)
    rc, its = g.solve_slab(x, prev, 0.0)
    assert rc == 0
    t = g.phase_timings(reset=True)
    # one preconditioner application per FGMRES step (the last one is not
    # counted as an iteration when it meets the tolerance)
    assert t["step"][1] == 1 and its <= t["gmg"][1] <= its + 1, (t, its)
    assert t["vmult"][1] > its and t["vanka"][1] > its
    assert 0 < t["gmg"][0] <= t["step"][0]
    assert g.phase_timings()["step"] == (0.0, 0)
    # the timed (eager) V-cycle gives the same solution as the default one
    h = st.Problem(dict(spec, print_timing=0))
    y = torch.zeros_like(x)
    rc2, its2 = h.solve_slab(y, prev, 0.0)
    assert rc2 == 0 and its2 == its and rel_err(y.cpu().numpy(), x.cpu().numpy()) < 1e-13
    g.close()
    h.close()
