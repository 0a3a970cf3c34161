"""The object-level C ABI composed the way tests/tp_01.cc:200-359 of
nlsmrg/dealii-stfem composes the reference's objects: level spaces,
SystemMatrix per level (operators.h:71-256), PreconditionVanka per level
(gmg.h:402-540), time and space two-level transfers (gmg.h:35-136), GMG
(gmg.h:652-906) and SolverFGMRES (time_integrators.h:44, :174), with the
caller's own time matrices from the time-FE entry points.  One slab of
TimeIntegratorHeat::solve (time_integrators.h:157-180) assembled from these
objects reproduces stfem_solve_slab bit for bit (the factored Vanka and the
two-pass operator are deterministic), and the objects outlive the handles
the caller drops."""
import numpy as np
import pytest

from helpers import box_vertices, log_uniform

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import stfem_amd as st  # noqa: E402

DEV = "cuda:0"
TAU, K, NONE = 1, 2, 3


def level_sequence(spec):
    """tp_01.cc:205-219: refinement, time degree, steps and step size of
    every level from the multigrid sequence of fe_time.h:43-69."""
    refine, r, nts, tau = spec["refine"], spec["r"], spec["nts"], spec["tau"]
    r_min = spec["r_min"]
    nts_min = max(1, nts // 2)
    seq = st.time_mg_sequence(refine + 1, r, r_min, nts, nts_min, False)
    L = len(seq)
    sref, rl, nl, tl = [0] * (L + 1), [0] * (L + 1), [0] * (L + 1), [0.0] * (L + 1)
    sref[L], rl[L], nl[L], tl[L] = refine, r, nts, tau
    for l in range(L - 1, -1, -1):
        t = seq[l]
        sref[l] = sref[l + 1] - (1 if t == NONE else 0)
        rl[l] = rl[l + 1] - (1 if t == K else 0)
        nl[l] = nl[l + 1] // (2 if t == TAU else 1)
        tl[l] = tl[l + 1] * (2.0 if t == TAU else 1.0)
    return seq, sref, rl, nl, tl


def build_objects(spec, X, kap, prec):
    ctx = st.Context()
    spaces = st.Space.hierarchy(ctx, spec, X, kap)
    seq, sref, rl, nl, tl = level_sequence(spec)
    type_ = spec["type"]
    ops, sms, trs = [], [], [None]
    for l in range(len(seq) + 1):
        A, B, _, _ = st.fe_time_weights(type_, rl[l], tl[l], nl[l])
        sp = spaces[sref[l]]
        ops.append(st.SystemMatrix(sp, A, B, prec))
        sms.append(st.PreconditionVanka(sp, A, B, prec, factored=True))
        if l > 0:
            t = seq[l - 1]
            if t == NONE:
                trs.append(st.Transfer.space(sp, spaces[sref[l - 1]], A.shape[0], prec))
            else:
                trs.append(st.Transfer.time(sp, type_, rl[l], nl[l], t, precision=prec))
    gmg = st.GMG(ops, sms, trs, options=dict(smoothing_steps=spec.get("smoothing_steps", 1)))
    A, B, G, Z = st.fe_time_weights(type_, spec["r"], spec["tau"], spec["nts"])
    fine = st.SystemMatrix(spaces[-1], A, B, 8)
    nt = A.shape[0]
    rK = G if type_ == st.CGP else np.zeros((nt, 1))
    rM = Z if type_ == st.CGP else G
    rhs_op = st.SystemMatrix(spaces[-1], rK, rM, 8)
    n = spaces[-1].info()["n_nodes"]
    solver = st.SolverFGMRES(ctx, n * nt)
    mask = torch.tensor(spaces[-1].constrained(), device=DEV)
    return dict(ctx=ctx, spaces=spaces, ops=ops, sms=sms, trs=trs, gmg=gmg, fine=fine,
                rhs_op=rhs_op, solver=solver, n=n, nt=nt, mask=mask)


def object_slab(o, prev):
    """TimeIntegratorHeat::solve from the objects (no source term)."""
    n, nt = o["n"], o["nt"]
    rhs = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    o["rhs_op"].vmult_add(rhs, prev)
    x = prev.repeat_interleave(nt).contiguous()  # extrapolate: every block = prev
    rc, its, _ = o["solver"].solve(o["fine"], x, rhs, P=o["gmg"])
    x.view(n, nt)[o["mask"]] = 0.0
    return rc, its, x


CASES = {
    "3d_q3_dg2_fp32": (dict(dim=3, refine=2, p=3, type=2, r=2, nts=2, r_min=0, tau=0.0625,
                            smoothing_steps=2), 0.2, True, 4),
    "3d_q3_dg2_fp64": (dict(dim=3, refine=2, p=3, type=2, r=2, nts=1, r_min=0, tau=0.125,
                            smoothing_steps=1), 0.2, True, 8),
    "2d_q2_cg2_fp64": (dict(dim=2, refine=3, p=2, type=1, r=2, nts=2, r_min=1, tau=0.03125),
                       0.1, False, 8),
}


@pytest.mark.parametrize("name", list(CASES))
def test_objects_reproduce_the_slab_solve(name):
    spec, perturb, kappa, prec = CASES[name]
    dim = spec["dim"]
    nc = [1 << spec["refine"]] * dim
    X = box_vertices(dim, nc, perturb=perturb, seed=7)
    kap = log_uniform(int(np.prod(nc)), 1e-2, 1e2, seed=11) if kappa else None
    full = dict(spec, source=1, vanka_factored=1, mg_float=1 if prec == 4 else 0)
    g = st.Problem(full, vertices=X, kappa=kap)
    o = build_objects(dict(spec, vanka_factored=1), X, kap, prec)
    assert o["gmg"].n_levels == g.n_levels
    for l in range(g.n_levels):
        assert o["gmg"].omega(l) == g.omega(l), l
    n, nt = g.n_nodes, g.nt
    assert (n, nt) == (o["n"], o["nt"])
    rng = np.random.default_rng(21)
    prev = torch.tensor(rng.standard_normal(n), dtype=torch.float64, device=DEV)
    prev[o["mask"]] = 0.0
    xg = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    print(f"[objects {name}] problem slab", flush=True)
    rc_g, its_g = g.solve_slab(xg, prev, 0.0)
    print(f"[objects {name}] object slab", flush=True)
    rc_o, its_o, xo = object_slab(o, prev)
    assert rc_g == 0 and rc_o == 0 and its_g == its_o, (rc_g, rc_o, its_g, its_o)
    assert torch.equal(xo, xg), (xo - xg).abs().max().item()
    # every level object applies what the problem's level applies
    dt = torch.float32 if prec == 4 else torch.float64
    for l in range(g.n_levels):
        info = g.level_info(l)
        x = torch.tensor(rng.standard_normal(info["n_nodes"] * info["nt"]), dtype=dt, device=DEV)
        y1, y2 = torch.zeros_like(x), torch.zeros_like(x)
        g.vmult(l, y1, x)
        o["ops"][l].vmult(y2, x)
        assert torch.equal(y1, y2), l
        g.vanka(l, y1, x)
        o["sms"][l].vmult(y2, x)
        assert torch.equal(y1, y2), l
    g.close()


def test_handles_may_be_dropped_in_any_order():
    spec, perturb, kappa, prec = CASES["3d_q3_dg2_fp64"]
    X = box_vertices(3, [4] * 3, perturb=perturb, seed=7)
    kap = log_uniform(64, 1e-2, 1e2, seed=11)
    o = build_objects(dict(spec, vanka_factored=1), X, kap, prec)
    n, nt = o["n"], o["nt"]
    x = torch.randn(n * nt, dtype=torch.float64, device=DEV)
    y1, y2 = torch.zeros_like(x), torch.zeros_like(x)
    o["gmg"].vmult(y1, x)
    # drop the caller's level, transfer, space and context handles
    for key in ("ops", "sms", "trs", "spaces"):
        for h in o[key]:
            if h is not None:
                h.close()
    o["ctx"].close()
    print("[objects] gmg after its parts were released", flush=True)
    o["gmg"].vmult(y2, x)
    assert torch.equal(y1, y2)
    rhs = torch.randn(n * nt, dtype=torch.float64, device=DEV)
    rhs.view(n, nt)[o["mask"]] = 0.0  # constrained rows of the operator are zero
    z = torch.zeros_like(rhs)
    rc, its, hist = o["solver"].solve(o["fine"], z, rhs, P=o["gmg"])
    assert rc == 0 and its > 0 and hist[-1] <= 1e-12 * hist[0] * 1.0001


def test_object_errors():
    ctx = st.Context()
    spaces = st.Space.hierarchy(ctx, dict(dim=2, refine=1, p=2))
    A = np.eye(2)
    with pytest.raises(st.StfemError):
        st.SystemMatrix(spaces[0], A, A, precision=2)
    with pytest.raises(st.StfemError):   # the same space twice: not one refinement apart
        st.Transfer.space(spaces[1], spaces[1], 2)
    op32 = st.SystemMatrix(spaces[1], A, A, 4)
    sm64 = st.PreconditionVanka(spaces[1], A, A, 8)
    with pytest.raises(st.StfemError):   # mixed precisions
        st.GMG([op32], [sm64], [])
    with pytest.raises(st.StfemError):   # a key that is not a multigrid option
        st.GMG([op32], [st.PreconditionVanka(spaces[1], A, A, 4)], [], options="refine=3")
    solver = st.SolverFGMRES(ctx, 10)
    with pytest.raises(st.StfemError):   # FP32 system operator
        solver.solve(op32, torch.zeros(10, dtype=torch.float64, device=DEV),
                     torch.zeros(10, dtype=torch.float64, device=DEV))
