"""GPU parity: the HIP path (libstfem_amd.so via the C ABI) against the CPU
oracle on the same seeded inputs.

Operators are compared at the level precision (FP64 to 1e-12 relative, FP32
to 2e-6); whole slab solves (to a 1e-10 relative residual) to 1e-12 relative
solution difference, with identical iteration counts in FP64 and within one
iteration with FP32 multigrid (measured: identical everywhere, differences
1e-15 to 1.1e-13).
"""
import numpy as np
import pytest

from helpers import blocks_to_nodes, box_vertices, log_uniform, nodes_to_blocks, rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
import stfem_amd as st  # noqa: E402

DEV = "cuda:0"

CASES = {
    # tp_01 tf01 shape: 2D DG(1), 2 steps at once, Q2, space+tau+k levels
    "2d_dg1_nts2": dict(spec=dict(dim=2, refine=2, type=2, r=1, nts=2, r_min=0), perturb=0.0,
                        kappa=False),
    # tp_01 tf02 shape: 2D CGP(2), 2 steps, FP32 multigrid
    "2d_cg2_nts2_fp32": dict(spec=dict(dim=2, refine=2, type=1, r=2, nts=2, r_min=1, mg_float=1),
                             perturb=0.0, kappa=False, oracle=dict(vanka_invert_T=0)),
    # distorted 2D mesh with a cell coefficient, DG(2), Q3
    "2d_dg2_distorted": dict(spec=dict(dim=2, refine=3, type=2, r=2, nts=1, p=3, r_min=0),
                             perturb=0.2, kappa=True),
    # the 32^3 benchmark's discretisation (Q3, dG(2), perturbed 0.2h, log-uniform kappa), 4^3 cells
    "3d_q3_dg2": dict(spec=dict(dim=3, refine=2, type=2, r=2, nts=1, p=3, r_min=0), perturb=0.2,
                      kappa=True),
    # the same with the factored Vanka inverse (eigen-decomposed patches)
    "3d_q3_dg2_factored": dict(spec=dict(dim=3, refine=2, type=2, r=2, nts=1, p=3, r_min=0,
                                         vanka_factored=1), perturb=0.2, kappa=True),
    # the 64^3 benchmark's discretisation (Q4, dG(3)), 2^3 cells, FP32 multigrid
    "3d_q4_dg3_fp32": dict(spec=dict(dim=3, refine=1, type=2, r=3, nts=1, p=4, r_min=2,
                                     mg_float=1), perturb=0.1, kappa=True,
                           oracle=dict(vanka_invert_T=0)),
    # the same with the factored Vanka set up on the device
    "3d_q4_dg3_fp32_factored": dict(spec=dict(dim=3, refine=1, type=2, r=3, nts=1, p=4, r_min=2,
                                              mg_float=1, vanka_factored=1),
                                    perturb=0.1, kappa=True),
    # FP32 multigrid with the factored Vanka (the benchmark configuration)
    "3d_q3_dg2_fp32_factored": dict(spec=dict(dim=3, refine=2, type=2, r=2, nts=1, p=3, r_min=0,
                                              mg_float=1, vanka_factored=1),
                                    perturb=0.2, kappa=True, oracle=dict(vanka_invert_T=0)),
    # the wave benchmark's time discretisation cG(2) with Q2 (heat operator), 4^3 cells
    "3d_q2_cg2": dict(spec=dict(dim=3, refine=2, type=1, r=2, nts=2, p=2, r_min=1), perturb=0.0,
                      kappa=False),
    # lowest degrees: Q1 space, dG(0) time with two steps per slab (tau levels
    # only), perturbed mesh and cell coefficient, factored Vanka, no Dirichlet
    # faces (with Q1 the one-cell coarse level would have no free node)
    "3d_q1_dg0_nts2": dict(spec=dict(dim=3, refine=2, type=2, r=0, nts=2, p=1, r_min=0,
                                     vanka_factored=1, dirichlet=0), perturb=0.2, kappa=True),
    # the same with Dirichlet faces: the one-cell levels have no free node
    # (relaxation parameter 1, smoother without effect there)
    "3d_q1_dg0_dirichlet": dict(spec=dict(dim=3, refine=2, type=2, r=0, nts=2, p=1, r_min=0,
                                          vanka_factored=1), perturb=0.2, kappa=True),
    # highest space degree: Q5 (216-node patches, the runtime-shape factored
    # kernel), dG(1), 2^3 cells, FP32 multigrid with device Vanka setup
    "3d_q5_dg1_fp32": dict(spec=dict(dim=3, refine=1, type=2, r=1, nts=1, p=5, r_min=0,
                                     mg_float=1, vanka_factored=1), perturb=0.1, kappa=True),
    # arithmetic-mean coarse coefficients, Q4 dG(3), FP32 factored Vanka
    "3d_q4_dg3_cc3": dict(spec=dict(dim=3, refine=2, type=2, r=3, nts=1, p=4, r_min=2,
                                    mg_float=1, vanka_factored=1, coarse_coef=3,
                                    smoothing_steps=2),
                          perturb=0.1, kappa=True),
}


def _build(name):
    c = CASES[name]
    spec = dict(c["spec"])
    dim = spec["dim"]
    nc = [1 << spec["refine"]] * dim
    X = box_vertices(dim, nc, perturb=c["perturb"], seed=7) if c["perturb"] > 0 else None
    kap = log_uniform(int(np.prod(nc)), 1e-2, 1e2, seed=11) if c["kappa"] else None
    # oracle-only keys (the explicit FP32 inverses formed in FP64 and then
    # rounded, as the library does) go to the oracle alone
    o = orc.Problem(dict(spec, **c.get("oracle", {})), coords=X, kappa=kap)
    g = st.Problem(spec, vertices=X, kappa=kap)
    return spec, o, g


_cache = {}


def problems(name):
    if name not in _cache:
        _cache.clear()
        _cache[name] = _build(name)
    return _cache[name]


def dtype_of(g):
    return torch.float32 if g.level_size == 4 else torch.float64


def tol_of(g):
    return 2e-6 if g.level_size == 4 else 1e-12


@pytest.mark.parametrize("name", list(CASES))
def test_hierarchy_matches(name):
    spec, o, g = problems(name)
    print(f"[case {name}] levels {g.n_levels}", flush=True)
    assert g.n_levels == o.n_levels
    for l in range(g.n_levels):
        a, b = g.level_info(l), o.level_info(l)
        for k in ("n_nodes", "nt", "r", "nts", "mg_type"):
            assert a[k] == b[k], (l, k, a, b)
    assert abs(g.tau - o.tau) < 1e-15


@pytest.mark.parametrize("name", list(CASES))
def test_level_operators(name):
    spec, o, g = problems(name)
    rng = np.random.default_rng(3)
    dt = dtype_of(g)
    for l in range(g.n_levels):
        info = g.level_info(l)
        n, nt = info["n_nodes"], info["nt"]
        x = rng.standard_normal(n * nt)  # includes Dirichlet entries: read as zero
        xg = torch.tensor(blocks_to_nodes(x, nt), dtype=dt, device=DEV)
        y = torch.zeros_like(xg)
        print(f"[case {name}] level {l} vmult n={n} nt={nt}", flush=True)
        g.vmult(l, y, xg)
        ref = o.level_apply(l, 0, x)
        assert rel_err(nodes_to_blocks(y.cpu().double().numpy(), nt), ref) < tol_of(g) * 10
        print(f"[case {name}] level {l} vanka", flush=True)
        g.vanka(l, y, xg)
        ref = o.level_apply(l, 1, x)
        assert rel_err(nodes_to_blocks(y.cpu().double().numpy(), nt), ref) < tol_of(g) * 100


@pytest.mark.parametrize("name", ["3d_q3_dg2", "3d_q4_dg3_fp32", "2d_dg1_nts2", "3d_q2_cg2",
                                  "2d_dg2_distorted"])
@pytest.mark.parametrize("chunking", [dict(op_split=0),
                                      dict(op_split=1 << 30),
                                      dict(op_kernel=0, op_split=0, op_jb=4, op_jb64=4),
                                      dict(op_kernel=0, op_split=0, op_jb=2, op_jb64=2),
                                      dict(op_kernel=0, op_split=0, op_jb=1, op_jb64=3),
                                      dict(op_kernel=0, op_split=1 << 30, op_jb=4, op_jb64=4)])
def test_operator_output_chunking(name, chunking):
    """Both 3D cell-pass kernels give the oracle's level and fine operators
    under every split of the output time components: the wavefront-group
    kernel (op_kernel=1, default) all components per workgroup, or one per
    grid row on grids below op_split workgroups; the workgroup kernel
    (op_kernel=0, also the 2D kernel) chunks of op_jb (FP32 levels) /
    op_jb64 (FP64) components side by side as grid rows."""
    spec, o, _ = problems(name)
    c = CASES[name]
    dim = spec["dim"]
    nc = [1 << spec["refine"]] * dim
    X = box_vertices(dim, nc, perturb=c["perturb"], seed=7) if c["perturb"] > 0 else None
    kap = log_uniform(int(np.prod(nc)), 1e-2, 1e2, seed=11) if c["kappa"] else None
    g = st.Problem(dict(spec, **chunking), vertices=X, kappa=kap)
    rng = np.random.default_rng(13)
    dt = dtype_of(g)
    for l in range(g.n_levels):
        info = g.level_info(l)
        n, nt = info["n_nodes"], info["nt"]
        x = rng.standard_normal(n * nt)
        xg = torch.tensor(blocks_to_nodes(x, nt), dtype=dt, device=DEV)
        y = torch.zeros_like(xg)
        print(f"[case {name}] {chunking} level {l} vmult", flush=True)
        g.vmult(l, y, xg)
        ref = o.level_apply(l, 0, x)
        assert rel_err(nodes_to_blocks(y.cpu().double().numpy(), nt), ref) < tol_of(g) * 10, l
    x = rng.standard_normal(g.n_nodes * g.nt)
    xg = torch.tensor(blocks_to_nodes(x, g.nt), dtype=torch.float64, device=DEV)
    y = torch.zeros_like(xg)
    for tr in (False, True):
        print(f"[case {name}] {chunking} fine vmult transpose={tr}", flush=True)
        (g.Tvmult if tr else g.vmult)(st.FINE, y, xg)
        assert rel_err(nodes_to_blocks(y.cpu().numpy(), g.nt), o.fine_vmult(x, transpose=tr)) < 1e-12
    g.close()


@pytest.mark.parametrize("name", list(CASES))
def test_transfers(name):
    spec, o, g = problems(name)
    rng = np.random.default_rng(5)
    dt = dtype_of(g)
    for l in range(1, g.n_levels):
        fi, ci = g.level_info(l), g.level_info(l - 1)
        xc = rng.standard_normal(ci["n_nodes"] * ci["nt"])
        yf0 = rng.standard_normal(fi["n_nodes"] * fi["nt"])
        print(f"[case {name}] prolongate to {l}", flush=True)
        dst = torch.tensor(blocks_to_nodes(yf0, fi["nt"]), dtype=dt, device=DEV)
        src = torch.tensor(blocks_to_nodes(xc, ci["nt"]), dtype=dt, device=DEV)
        g.prolongate_and_add(l, dst, src)
        ref = o.level_apply(l, 2, xc, dst=yf0.copy())
        assert rel_err(nodes_to_blocks(dst.cpu().double().numpy(), fi["nt"]), ref) < tol_of(g) * 10
        xf = rng.standard_normal(fi["n_nodes"] * fi["nt"])
        yc0 = rng.standard_normal(ci["n_nodes"] * ci["nt"])
        print(f"[case {name}] restrict from {l}", flush=True)
        dst = torch.tensor(blocks_to_nodes(yc0, ci["nt"]), dtype=dt, device=DEV)
        src = torch.tensor(blocks_to_nodes(xf, fi["nt"]), dtype=dt, device=DEV)
        g.restrict_and_add(l, dst, src)
        ref = o.level_apply(l, 3, xf, dst=yc0.copy())
        assert rel_err(nodes_to_blocks(dst.cpu().double().numpy(), ci["nt"]), ref) < tol_of(g) * 10


@pytest.mark.parametrize("name", list(CASES))
def test_relaxation_parameters(name):
    spec, o, g = problems(name)
    for l in range(g.n_levels):
        a, b = g.omega(l), o.omega(l)
        tol = 1e-4 if g.level_size == 4 else 1e-8
        assert abs(a - b) <= tol * abs(b), (l, a, b)


@pytest.mark.parametrize("name", list(CASES))
def test_fine_vmult_and_transpose(name):
    spec, o, g = problems(name)
    rng = np.random.default_rng(9)
    n, nt = g.n_nodes, g.nt
    x = rng.standard_normal(n * nt)
    xg = torch.tensor(blocks_to_nodes(x, nt), dtype=torch.float64, device=DEV)
    y = torch.zeros_like(xg)
    for tr in (False, True):
        print(f"[case {name}] fine vmult transpose={tr}", flush=True)
        (g.Tvmult if tr else g.vmult)(st.FINE, y, xg)
        ref = o.fine_vmult(x, transpose=tr)
        assert rel_err(nodes_to_blocks(y.cpu().numpy(), nt), ref) < 1e-12


@pytest.mark.parametrize("name", list(CASES))
def test_gmg_vcycle(name):
    spec, o, g = problems(name)
    rng = np.random.default_rng(13)
    n, nt = g.n_nodes, g.nt
    x = rng.standard_normal(n * nt)
    xg = torch.tensor(blocks_to_nodes(x, nt), dtype=torch.float64, device=DEV)
    y = torch.zeros_like(xg)
    print(f"[case {name}] gmg vmult", flush=True)
    g.gmg_vmult(y, xg)
    ref = o.gmg_vmult(x)
    tol = 1e-4 if g.level_size == 4 else 1e-10
    assert rel_err(nodes_to_blocks(y.cpu().numpy(), nt), ref) < tol


@pytest.mark.parametrize("name", list(CASES))
def test_slab_solves(name):
    spec, o, g = problems(name)
    n, nt = g.n_nodes, g.nt
    xo = np.zeros(n * nt)
    xg = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    prev_g = torch.zeros(n, dtype=torch.float64, device=DEV)
    time = 0.0
    for k in range(3):
        prev_o = xo[(nt - 1) * n:].copy()
        prev_g.copy_(torch.tensor(prev_o, device=DEV))
        rc_o, xo, its_o = orc.solve_slab(o, xo, prev_o, time)
        print(f"[case {name}] slab {k}", flush=True)
        rc_g, its_g = g.solve_slab(xg, prev_g, time)
        assert rc_o == 0 and rc_g == 0
        if g.level_size == 8:
            assert its_g == its_o, (k, its_g, its_o)
        else:
            assert abs(its_g - its_o) <= 1, (k, its_g, its_o)
        err = rel_err(nodes_to_blocks(xg.cpu().numpy(), nt), xo)
        print(f"[case {name}] slab {k}: its {its_g} (oracle {its_o}), rel diff {err:.2e}",
              flush=True)
        assert err < 1e-12, (k, err)
        time += spec["nts"] * g.tau


@pytest.mark.parametrize("name", ["3d_q3_dg2", "2d_dg1_nts2", "3d_q2_cg2", "2d_dg2_distorted",
                                  "3d_q3_dg2_fp32_factored", "3d_q4_dg3_fp32"])
@pytest.mark.parametrize("ortho", [0, 1])
def test_fgmres_residual_history(name, ortho):
    """FGMRES with the GMG preconditioner from x = 0 on b = A x_r (x_r
    seeded): with either Gram-Schmidt form (0: chain of axpy+dot passes, 1:
    one pass of dots corrected by the Gram matrix) the GPU solver takes the
    oracle's iterations and its residual history (SolverFGMRES's projected
    residuals) agrees within 1e-10 of the initial residual (FP32 multigrid:
    within one iteration and 1e-6 of it); the solutions agree to 1e-12."""
    spec, o, _ = problems(name)
    c = CASES[name]
    dim = spec["dim"]
    nc = [1 << spec["refine"]] * dim
    X = box_vertices(dim, nc, perturb=c["perturb"], seed=7) if c["perturb"] > 0 else None
    kap = log_uniform(int(np.prod(nc)), 1e-2, 1e2, seed=11) if c["kappa"] else None
    g = st.Problem(dict(spec, orthogonalization=ortho), vertices=X, kappa=kap)
    rng = np.random.default_rng(29)
    n, nt = g.n_nodes, g.nt
    b = o.fine_vmult(rng.standard_normal(n * nt))
    rc_o, xo, its_o, ho = o.fgmres(b, np.zeros(n * nt))
    xg = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    bg = torch.tensor(blocks_to_nodes(b, nt), dtype=torch.float64, device=DEV)
    print(f"[case {name}] fgmres orthogonalization {ortho}", flush=True)
    rc_g, its_g, hg = g.fgmres(xg, bg)
    assert rc_o == 0 and rc_g == 0, (rc_o, rc_g)
    fp64 = g.level_size == 8
    if fp64:
        assert its_g == its_o, (its_g, its_o)
    else:
        assert abs(its_g - its_o) <= 1, (its_g, its_o)
    k = min(len(hg), len(ho))
    dev = np.abs(hg[:k] - ho[:k]).max() / ho[0]
    assert dev < (1e-10 if fp64 else 1e-6), dev
    err = rel_err(nodes_to_blocks(xg.cpu().numpy(), nt), xo)
    print(f"[case {name}] its {its_g} (oracle {its_o}), history dev {dev:.2e}, "
          f"rel diff {err:.2e}", flush=True)
    assert err < 1e-12, err
    g.close()


@pytest.mark.parametrize("name", ["3d_q3_dg2_factored", "3d_q3_dg2_fp32_factored",
                                  "3d_q4_dg3_fp32_factored", "3d_q4_dg3_cc3"])
def test_fused_residual_sweeps(name):
    """Relaxation sweeps that sum A u at the patch nodes from the operator's
    cell buffer (on levels of up to fuse_residual cells) give the same
    V-cycle, to the last bit, as sweeps through the operator's node gather
    (fuse_residual=0), and the same slab solve."""
    spec, o, _ = problems(name)
    c = CASES[name]
    dim = spec["dim"]
    nc = [1 << spec["refine"]] * dim
    X = box_vertices(dim, nc, perturb=c["perturb"], seed=7) if c["perturb"] > 0 else None
    kap = log_uniform(int(np.prod(nc)), 1e-2, 1e2, seed=11) if c["kappa"] else None
    a = st.Problem(dict(spec, fuse_residual=1 << 30, smoothing_steps=2), vertices=X, kappa=kap)
    b = st.Problem(dict(spec, fuse_residual=0, smoothing_steps=2), vertices=X, kappa=kap)
    rng = np.random.default_rng(37)
    n, nt = a.n_nodes, a.nt
    x = torch.tensor(rng.standard_normal(n * nt), dtype=torch.float64, device=DEV)
    ya, yb = torch.zeros_like(x), torch.zeros_like(x)
    print(f"[case {name}] fused residual: gmg vmult", flush=True)
    a.gmg_vmult(ya, x)
    b.gmg_vmult(yb, x)
    assert torch.equal(ya, yb), (ya - yb).abs().max().item()
    xa, xb = torch.zeros_like(x), torch.zeros_like(x)
    prev = torch.zeros(n, dtype=torch.float64, device=DEV)
    print(f"[case {name}] fused residual: slab solve", flush=True)
    ra, ia = a.solve_slab(xa, prev, 0.0)
    rb, ib = b.solve_slab(xb, prev, 0.0)
    assert ra == 0 and rb == 0 and ia == ib, (ra, rb, ia, ib)
    assert torch.equal(xa, xb), (xa - xb).abs().max().item()
    a.close()
    b.close()


def test_transfer_timing_and_bytes_model():
    """The measurement entry points for the transfers between the two finest
    levels: algorithmic bytes (source read, destination read and written)
    and HIP-event timings."""
    spec, o, g = problems("3d_q3_dg2")
    L = g.n_levels - 1
    fi, ci = g.level_info(L), g.level_info(L - 1)
    nf, nc, s = fi["n_nodes"] * fi["nt"], ci["n_nodes"] * ci["nt"], g.level_size
    assert g.bytes_model(2, L) == s * (nc + 2 * nf)
    assert g.bytes_model(3, L) == s * (nf + 2 * nc)
    print("[transfer timing]", flush=True)
    assert g.time_kernel(4, L, warmup=1, reps=2) > 0
    assert g.time_kernel(5, L, warmup=1, reps=2) > 0


@pytest.mark.parametrize("name", ["3d_q3_dg2", "3d_q4_dg3_fp32_factored"])
def test_node_gather_unaligned_vectors(name):
    """The node gathers access a node's time components as vectors when the
    node vectors are 16-byte aligned and fall back to scalar accesses
    otherwise: a destination offset by one value gives the same bits (the
    operator, and the factored Vanka's two-pass form)."""
    spec, o, g = problems(name)
    rng = np.random.default_rng(29)
    dt = dtype_of(g)
    for l in (g.n_levels - 1, max(0, g.n_levels - 2)):
        info = g.level_info(l)
        m = info["n_nodes"] * info["nt"]
        x = torch.tensor(rng.standard_normal(m), dtype=dt, device=DEV)
        y = torch.zeros(m, dtype=dt, device=DEV)
        buf = torch.zeros(m + 1, dtype=dt, device=DEV)
        print(f"[case {name}] level {l} unaligned gather", flush=True)
        g.vmult(l, y, x)
        g.vmult(l, buf[1:], x)
        assert buf[1:].data_ptr() % 16 != 0
        assert torch.equal(y, buf[1:]), l
        g.vanka(l, y, x)
        g.vanka(l, buf[1:], x)
        if spec.get("vanka_factored"):
            assert torch.equal(y, buf[1:]), l
        else:
            # explicit inverses add their corrections with atomics: equal
            # up to the order of summation
            assert rel_err(buf[1:].cpu().numpy(), y.cpu().numpy()) < tol_of(g), l


@pytest.mark.parametrize("name", ["3d_q3_dg2", "3d_q4_dg3_fp32_factored"])
def test_transfers_unaligned_vectors(name):
    """The space transfer sweeps take one thread per node with vector
    accesses when the vectors are 16-byte aligned and one thread per entry
    otherwise; both sum a row's taps in the same order, so source and
    destination vectors offset by one value give the same bits."""
    spec, o, g = problems(name)
    rng = np.random.default_rng(41)
    dt = dtype_of(g)
    for l in range(1, g.n_levels):
        fi, ci = g.level_info(l), g.level_info(l - 1)
        mf, mc = fi["n_nodes"] * fi["nt"], ci["n_nodes"] * ci["nt"]
        for what, (m_src, m_dst) in (("prolongate", (mc, mf)), ("restrict", (mf, mc))):
            x = torch.tensor(rng.standard_normal(m_src), dtype=dt, device=DEV)
            y0 = torch.tensor(rng.standard_normal(m_dst), dtype=dt, device=DEV)
            xs = torch.zeros(m_src + 1, dtype=dt, device=DEV)
            ys = torch.zeros(m_dst + 1, dtype=dt, device=DEV)
            xs[1:] = x
            ys[1:] = y0
            y = y0.clone()
            assert xs[1:].data_ptr() % 16 != 0 and ys[1:].data_ptr() % 16 != 0
            print(f"[case {name}] level {l} unaligned {what}", flush=True)
            f = g.prolongate_and_add if what == "prolongate" else g.restrict_and_add
            f(l, y, x)
            f(l, ys[1:], xs[1:])
            assert torch.equal(y, ys[1:]), (l, what)


@pytest.mark.parametrize("name", ["3d_q3_dg2_factored", "3d_q4_dg3_fp32_factored"])
def test_gmg_vmult_unaligned_vectors(name):
    """The V-cycle's precision conversions (FP64 vectors in, level precision
    inside) use 16-byte accesses on aligned vectors and scalar ones
    otherwise: FP64 input and output offset by one value give the same
    bits (factored Vanka: the V-cycle has no atomics, so two applications
    agree bit for bit)."""
    spec, o, g = problems(name)
    rng = np.random.default_rng(43)
    m = g.n_nodes * g.nt
    x = torch.tensor(rng.standard_normal(m), dtype=torch.float64, device=DEV)
    xs = torch.zeros(m + 1, dtype=torch.float64, device=DEV)
    xs[1:] = x
    y = torch.zeros(m, dtype=torch.float64, device=DEV)
    ys = torch.zeros(m + 1, dtype=torch.float64, device=DEV)
    assert xs[1:].data_ptr() % 16 != 0 and ys[1:].data_ptr() % 16 != 0
    print(f"[case {name}] unaligned gmg vmult", flush=True)
    g.gmg_vmult(y, x)
    g.gmg_vmult(ys[1:], xs[1:])
    assert torch.equal(y, ys[1:])


def test_vanka_sweep_kernel_timing():
    """The sweep kernel of the factored Vanka can be timed alone (selector 6:
    the cell pass without the node gather), and a level whose Vanka has no
    two-pass buffer (explicit inverses) refuses it with an error status."""
    spec, o, g = problems("3d_q3_dg2_fp32_factored")
    L = g.n_levels - 1
    print("[vanka sweep kernel timing]", flush=True)
    assert g.time_kernel(6, L, warmup=1, reps=2) > 0
    spec, o, g = problems("3d_q3_dg2")  # explicit inverses
    with pytest.raises(RuntimeError):
        g.time_kernel(6, g.n_levels - 1, warmup=1, reps=2)


def test_graph_replay_matches_eager():
    """The V-cycle replayed from a captured HIP graph equals the eagerly
    launched V-cycle on every replay, with other device work in between."""
    spec, o, g = problems("3d_q3_dg2")
    c = CASES["3d_q3_dg2"]
    dim = spec["dim"]
    nc = [1 << spec["refine"]] * dim
    X = box_vertices(dim, nc, perturb=c["perturb"], seed=7)
    kap = log_uniform(int(np.prod(nc)), 1e-2, 1e2, seed=11)
    ge = st.Problem(dict(spec, use_graph=0), vertices=X, kappa=kap)
    rng = np.random.default_rng(17)
    n, nt = g.n_nodes, g.nt
    for rep in range(4):
        x = torch.tensor(rng.standard_normal(n * nt), dtype=torch.float64, device=DEV)
        y1, y2, tmp = torch.zeros_like(x), torch.zeros_like(x), torch.zeros_like(x)
        print(f"[graph replay {rep}]", flush=True)
        g.gmg_vmult(y1, x)          # graph replay (captured on first use)
        g.vmult(st.FINE, tmp, x)    # eager work between replays
        ge.gmg_vmult(y2, x)         # eager V-cycle
        assert rel_err(y1.cpu().numpy(), y2.cpu().numpy()) < 1e-13, rep
    ge.close()


DENSE_CASES = ["3d_q3_dg2", "3d_q3_dg2_fp32_factored", "3d_q4_dg3_fp32", "2d_dg1_nts2",
               "3d_q2_cg2", "2d_cg2_nts2_fp32"]


@pytest.mark.parametrize("dense", [0, 200])
@pytest.mark.parametrize("name", DENSE_CASES)
def test_dense_coarse_operator(name, dense):
    """The V-cycle whose coarse part is one precomputed matrix (default: the
    largest level below the finest with at most 8192 unknowns) equals the
    V-cycle launched level by level (dense_coarse_dofs=0) and one collapsed
    at a smaller level (200 unknowns); both match the oracle."""
    spec, o, g = problems(name)
    c = CASES[name]
    dim = spec["dim"]
    nc = [1 << spec["refine"]] * dim
    X = box_vertices(dim, nc, perturb=c["perturb"], seed=7) if c["perturb"] > 0 else None
    kap = log_uniform(int(np.prod(nc)), 1e-2, 1e2, seed=11) if c["kappa"] else None
    h = st.Problem(dict(spec, dense_coarse_dofs=dense), vertices=X, kappa=kap)
    full = [g.level_info(l) for l in range(g.n_levels)]
    part = [h.level_info(l) for l in range(h.n_levels)]
    L = g.n_levels - 1
    assert full[L]["in_dense"] == 0 and full[L - 1]["in_dense"] == 1, full
    for i in part:
        m = i["n_nodes"] * i["nt"]
        if dense == 0:
            assert i["in_dense"] == 0, part
        elif i["in_dense"]:
            assert m <= dense
    rng = np.random.default_rng(23)
    n, nt = g.n_nodes, g.nt
    x = rng.standard_normal(n * nt)
    xg = torch.tensor(blocks_to_nodes(x, nt), dtype=torch.float64, device=DEV)
    y1, y2 = torch.zeros_like(xg), torch.zeros_like(xg)
    print(f"[case {name}] dense_coarse_dofs {dense}: gmg vmult", flush=True)
    g.gmg_vmult(y1, xg)
    h.gmg_vmult(y2, xg)
    fp32 = g.level_size == 4
    assert rel_err(y1.cpu().numpy(), y2.cpu().numpy()) < (1e-4 if fp32 else 1e-10)
    ref = o.gmg_vmult(x)
    assert rel_err(nodes_to_blocks(y2.cpu().numpy(), nt), ref) < (1e-4 if fp32 else 1e-10)
    h.close()


@pytest.mark.parametrize("name,method", [
    ("3d_q3_dg2_factored", 3), ("3d_q3_dg2_fp32_factored", 3), ("3d_q3_dg2_factored", 1),
    ("3d_q3_dg2_factored", 2), ("3d_q3_dg2_factored", 4), ("3d_q3_dg2_factored", 5),
    ("3d_q3_dg2_factored", 6), ("3d_q4_dg3_fp32_factored", 3), ("3d_q4_dg3_fp32_factored", 4)])
def test_vanka_device_setup_matches_host(name, method):
    """The factored Vanka set up on the GPU (element and patch matrices,
    batched generalized eigensolver, time blocks) applies the same operator
    as the one set up on the host, on every level, for each eigensolver:
    1 Jacobi, 2 divide and conquer, 4/5 Cholesky reduction + divide and
    conquer/Jacobi, 6 the hand-written batched eigensolver, 3 the automatic
    choice (6 for patches of up to 128 nodes)."""
    spec, o, _ = problems(name)
    c = CASES[name]
    g = st.Problem(dict(spec, vanka_device_setup=method),
                   vertices=box_vertices(spec["dim"], [1 << spec["refine"]] * spec["dim"],
                                         perturb=c["perturb"], seed=7),
                   kappa=log_uniform((1 << spec["refine"]) ** spec["dim"], 1e-2, 1e2, seed=11))
    dim = spec["dim"]
    nc = [1 << spec["refine"]] * dim
    X = box_vertices(dim, nc, perturb=c["perturb"], seed=7)
    kap = log_uniform(int(np.prod(nc)), 1e-2, 1e2, seed=11)
    h = st.Problem(dict(spec, vanka_device_setup=0), vertices=X, kappa=kap)
    rng = np.random.default_rng(31)
    dt = dtype_of(g)
    for l in range(g.n_levels):
        info = g.level_info(l)
        x = torch.tensor(rng.standard_normal(info["n_nodes"] * info["nt"]), dtype=dt, device=DEV)
        y1, y2 = torch.zeros_like(x), torch.zeros_like(x)
        print(f"[case {name}] level {l} vanka device setup {method} vs host", flush=True)
        g.vanka(l, y1, x)
        h.vanka(l, y2, x)
        tol = 1e-5 if g.level_size == 4 else 1e-11
        assert rel_err(y1.cpu().double().numpy(), y2.cpu().double().numpy()) < tol, l
    h.close()
    g.close()


def test_single_cell_levels_shortcut():
    """s smoothing sweeps on a one-cell level (where the Vanka patch is the
    whole level) equal one sweep with 1-(1-omega)^s."""
    spec = dict(dim=2, refine=2, type=2, r=2, nts=2, r_min=0)
    a = st.Problem(dict(spec, exact_single_cell=1))
    b = st.Problem(dict(spec, exact_single_cell=0))
    rng = np.random.default_rng(19)
    x = torch.tensor(rng.standard_normal(a.n_nodes * a.nt), dtype=torch.float64, device=DEV)
    ya, yb = torch.zeros_like(x), torch.zeros_like(x)
    a.gmg_vmult(ya, x)
    b.gmg_vmult(yb, x)
    assert rel_err(ya.cpu().numpy(), yb.cpu().numpy()) < 1e-12


@pytest.mark.parametrize("name", ["3d_q3_dg2_fp32_factored", "3d_q4_dg3_fp32_factored",
                                  "3d_q3_dg2_factored"])
@pytest.mark.parametrize("variant", [2, 3, 4, 5])
def test_vanka_kernel_variants(name, variant):
    """The factored sweep kernels (2 LDS-staged S, 3 register-tiled S, 4 its
    software-pipelined persistent form in FP32, 5 the FP32 form streaming S
    through LDS with direct global-to-LDS loads) on every level against the
    oracle's PreconditionVanka::vmult."""
    spec, o, _ = problems(name)
    c = CASES[name]
    dim = spec["dim"]
    nc = [1 << spec["refine"]] * dim
    X = box_vertices(dim, nc, perturb=c["perturb"], seed=7) if c["perturb"] > 0 else None
    kap = log_uniform(int(np.prod(nc)), 1e-2, 1e2, seed=11) if c["kappa"] else None
    g = st.Problem(dict(spec, vanka_kernel=variant), vertices=X, kappa=kap)
    rng = np.random.default_rng(5)
    dt = dtype_of(g)
    for l in range(g.n_levels):
        info = g.level_info(l)
        n, nt = info["n_nodes"], info["nt"]
        x = rng.standard_normal(n * nt)
        xg = torch.tensor(blocks_to_nodes(x, nt), dtype=dt, device=DEV)
        y = torch.zeros_like(xg)
        print(f"[case {name}] variant {variant} level {l} vanka", flush=True)
        g.vanka(l, y, xg)
        ref = o.level_apply(l, 1, x)
        assert rel_err(nodes_to_blocks(y.cpu().double().numpy(), nt), ref) < tol_of(g) * 100
    g.close()


def test_streamed_sweep_over_many_cells_per_workgroup():
    """The LDS-streamed FP32 sweep (vanka_kernel=5) runs one persistent
    workgroup per CU over a range of cells, loading the next cell's S while
    it works on the current one.  On 8^3 cells (two cells per workgroup on
    256 CUs, Q3/dG(2)) it matches the oracle's PreconditionVanka::vmult on
    the finest level; on 16^3 cells (16 cells per workgroup, Q4/dG(3)) it
    matches the register-tiled sweep (pinned to the oracle above) on every
    level, including the two-cell rounding at the ends of the ranges."""
    rng = np.random.default_rng(44)
    # Q3/dG(2), 8^3 cells, against the oracle (finest level)
    spec = dict(dim=3, refine=3, type=2, r=2, nts=1, p=3, r_min=0, mg_float=1, vanka_factored=1,
                mg_space=0, space_time_mg=0)
    nc = [8] * 3
    X = box_vertices(3, nc, perturb=0.2, seed=7)
    kap = log_uniform(512, 1e-2, 1e2, seed=11)
    o = orc.Problem(dict(spec, vanka_invert_T=0), coords=X, kappa=kap)
    g = st.Problem(dict(spec, vanka_kernel=5), vertices=X, kappa=kap)
    L = g.n_levels - 1
    info = g.level_info(L)
    n, nt = info["n_nodes"], info["nt"]
    x = rng.standard_normal(n * nt)
    xg = torch.tensor(blocks_to_nodes(x, nt), dtype=torch.float32, device=DEV)
    y = torch.zeros_like(xg)
    print("[streamed sweep] 8^3 Q3 dG(2) vs oracle", flush=True)
    g.vanka(L, y, xg)
    ref = o.level_apply(L, 1, x)
    assert rel_err(nodes_to_blocks(y.cpu().double().numpy(), nt), ref) < 2e-4
    g.close()
    # Q4/dG(3), 16^3 cells, against the register-tiled sweep
    spec = dict(dim=3, refine=4, type=2, r=3, nts=1, p=4, r_min=2, mg_float=1, vanka_factored=1,
                coarse_coef=3, dense_coarse_dofs=0)
    nc = [16] * 3
    X = box_vertices(3, nc, perturb=0.1, seed=2026)
    kap = log_uniform(4096, 1e-3, 1e3, seed=2027)
    a = st.Problem(dict(spec, vanka_kernel=5), vertices=X, kappa=kap)
    b = st.Problem(dict(spec, vanka_kernel=3), vertices=X, kappa=kap)
    for l in range(a.n_levels):
        info = a.level_info(l)
        xl = torch.tensor(rng.standard_normal(info["n_nodes"] * info["nt"]), dtype=torch.float32,
                          device=DEV)
        ya, yb = torch.zeros_like(xl), torch.zeros_like(xl)
        print(f"[streamed sweep] 16^3 Q4 dG(3) level {l}", flush=True)
        a.vanka(l, ya, xl)
        b.vanka(l, yb, xl)
        assert rel_err(ya.cpu().double().numpy(), yb.cpu().double().numpy()) < 1e-5, l
    # and the whole slab solve takes the same iterations
    n, nt = a.n_nodes, a.nt
    xa = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    xb = torch.zeros_like(xa)
    prev = torch.zeros(n, dtype=torch.float64, device=DEV)
    print("[streamed sweep] 16^3 slab solve", flush=True)
    ra, ia = a.solve_slab(xa, prev, 0.0)
    rb, ib = b.solve_slab(xb, prev, 0.0)
    assert ra == 0 and rb == 0 and abs(ia - ib) <= 1, (ia, ib)
    assert rel_err(xa.cpu().numpy(), xb.cpu().numpy()) < 1e-10
    a.close()
    b.close()


@pytest.mark.parametrize("dim,p,refine", [(3, 4, 1), (3, 2, 2), (2, 5, 2)])
def test_own_eigensolver_on_symmetric_patches(dim, p, refine):
    """The hand-written batched eigensolver (Cholesky, Householder, implicit
    QL with a rotation log) on an unperturbed mesh with a constant
    coefficient, where the patch matrices have the symmetries of the square
    or cube and hence exactly repeated eigenvalues: the Vanka it sets up
    matches the host setup (dense generalized eigensolver) on every level, in
    FP64, and the slab solve takes the same iterations."""
    spec = dict(dim=dim, refine=refine, type=2, r=1, nts=1, p=p, r_min=0, vanka_factored=1)
    a = st.Problem(dict(spec, vanka_device_setup=6))
    b = st.Problem(dict(spec, vanka_device_setup=0))
    rng = np.random.default_rng(5)
    for l in range(a.n_levels):
        info = a.level_info(l)
        x = torch.tensor(rng.standard_normal(info["n_nodes"] * info["nt"]), dtype=torch.float64,
                         device=DEV)
        ya, yb = torch.zeros_like(x), torch.zeros_like(x)
        print(f"[own eigensolver] dim {dim} Q{p} level {l}", flush=True)
        a.vanka(l, ya, x)
        b.vanka(l, yb, x)
        assert rel_err(ya.cpu().numpy(), yb.cpu().numpy()) < 1e-11, l
    n, nt = a.n_nodes, a.nt
    xa = torch.zeros(n * nt, dtype=torch.float64, device=DEV)
    xb = torch.zeros_like(xa)
    prev = torch.ones(n, dtype=torch.float64, device=DEV)
    ra, ia = a.solve_slab(xa, prev, 0.0)
    rb, ib = b.solve_slab(xb, prev, 0.0)
    assert ra == 0 and rb == 0 and ia == ib, (ia, ib)
    assert rel_err(xa.cpu().numpy(), xb.cpu().numpy()) < 1e-10
    a.close()
    b.close()
