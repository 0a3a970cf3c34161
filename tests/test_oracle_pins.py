"""The oracle's multigrid components pinned directly against dense linear
algebra and exact polynomial reproduction, independently of the golden
iteration counts.

* PreconditionVanka (tests/include/gmg.h:409-514 of nlsmrg/dealii-stfem):
  for every cell c the block B_c = Alpha (x) K[I_c, I_c] + Beta (x) M[I_c, I_c]
  of the assembled level matrices (time-major patch index k + i n_c,
  gmg.h:460-468), inverted densely with numpy, applied to the gathered
  patch vector and scattered with weight 1/valence (valence = number of
  cells that hold the dof, gmg.h:424-441).
* The space transfer (MGTwoLevelBlockTransfer, gmg.h:35-75, built on deal.II's
  MGTwoLevelTransfer at gmg.h:371-380): the prolongation is the FE embedding,
  so it maps the coarse interpolant of any Q_p function of the logical
  coordinates to its fine interpolant exactly (zero Dirichlet values where
  the function vanishes on the boundary); the restriction is its transpose
  (restrict_is_transpose_prolongate, gmg.h:557), Dirichlet rows included.
* The relaxation estimate (GMG::estimate_relaxation, gmg.h:792-844): the
  power-iteration estimate of the largest eigenvalue of V A (V the Vanka
  operator, A the level operator) against numpy's dense eigenvalues, and
  omega = 2 / (alpha + lambda) with lambda = 1.15 * estimate and
  alpha = 0.9 * lambda (power iteration, smoothing range 1).
"""
import numpy as np
import pytest

import oracle as orc
from helpers import box_vertices, log_uniform


def _problem(spec, perturb=0.0, kappa=False, seed=7):
    dim = spec["dim"]
    nc = [1 << spec["refine"]] * dim
    X = box_vertices(dim, nc, perturb=perturb, seed=seed) if perturb > 0 else None
    kap = log_uniform(int(np.prod(nc)), 1e-2, 1e2, seed=11) if kappa else None
    return orc.Problem(spec, coords=X, kappa=kap)


def cell_dofs(info, p, dim):
    """Lexicographic node numbers of every cell's (p+1)^dim nodes (x fastest)."""
    nc = [info["nc0"], info["nc1"], info["nc2"]][:dim]
    nn = [c * p + 1 for c in nc]
    loc = np.array(list(np.ndindex(*([p + 1] * dim))))[:, ::-1]  # x fastest
    out = []
    for cell in np.ndindex(*nc[::-1]):
        c = cell[::-1]
        g = [c[d] * p + loc[:, d] for d in range(dim)]
        idx = g[0] + nn[0] * g[1]
        if dim == 3:
            idx = idx + nn[0] * nn[1] * g[2]
        out.append(idx)
    return out


def brute_force_vanka(K, M, Alpha, Beta, cells, x):
    """dst = sum_c (1/valence) B_c^{-1} x|_c with block-major x (nt blocks)."""
    n = K.shape[0]
    nt = Alpha.shape[0]
    valence = np.zeros(n)
    for I in cells:
        valence[I] += 1.0
    X = x.reshape(nt, n)
    dst = np.zeros((nt, n))
    for I in cells:
        Kc, Mc = K[np.ix_(I, I)], M[np.ix_(I, I)]
        B = np.kron(Alpha, Kc) + np.kron(Beta, Mc)  # row index i * n_c + k (time-major)
        loc = np.linalg.solve(B, X[:, I].reshape(-1)).reshape(nt, len(I))
        dst[:, I] += loc / valence[I]
    return dst.reshape(-1)


VANKA_CASES = {
    "2d_q2_dg1_nts2": (dict(dim=2, refine=2, p=2, type=2, r=1, nts=2, r_min=0), 0.0, False),
    "2d_q3_dg2_perturbed": (dict(dim=2, refine=2, p=3, type=2, r=2, nts=1, r_min=0), 0.2, True),
    "2d_q2_cg2": (dict(dim=2, refine=2, p=2, type=1, r=2, nts=2, r_min=1), 0.1, True),
    "3d_q2_dg1": (dict(dim=3, refine=1, p=2, type=2, r=1, nts=1, r_min=0), 0.2, True),
    "3d_q3_dg2_no_dirichlet": (dict(dim=3, refine=1, p=3, type=2, r=2, nts=1, r_min=0,
                                    dirichlet=0), 0.15, True),
    "3d_q2_cg2_wave": (dict(dim=3, refine=1, p=2, type=1, r=2, nts=1, r_min=1, problem=1), 0.0,
                       False),
}


@pytest.mark.parametrize("name", list(VANKA_CASES))
def test_vanka_equals_dense_patch_inverses(name):
    spec, perturb, kappa = VANKA_CASES[name]
    P = _problem(spec, perturb, kappa)
    rng = np.random.default_rng(41)
    checked = 0
    for l in range(P.n_levels):
        info = P.level_info(l)
        if info["n_nodes"] > 1500:
            continue
        K, M = P.level_matrix(l, 0), P.level_matrix(l, 1)
        A, B = P.level_time_matrices(l)
        cells = cell_dofs(info, spec["p"], spec["dim"])
        assert len(cells) == info["n_cells"]
        x = rng.standard_normal(info["n_nodes"] * info["nt"])
        ref = brute_force_vanka(K, M, A, B, cells, x)
        got = P.level_apply(l, 1, x)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < 1e-12, (l, err)
        checked += 1
    assert checked >= 2


def test_vanka_pin_detects_wrong_index_order():
    """The brute force above fails for a plausible mistake: the patch block
    built space-major (k n_t + i) instead of time-major."""
    spec, perturb, kappa = VANKA_CASES["2d_q2_dg1_nts2"]
    P = _problem(spec, perturb, kappa)
    l = P.n_levels - 1
    info = P.level_info(l)
    K, M = P.level_matrix(l, 0), P.level_matrix(l, 1)
    A, B = P.level_time_matrices(l)
    cells = cell_dofs(info, spec["p"], spec["dim"])
    x = np.random.default_rng(3).standard_normal(info["n_nodes"] * info["nt"])
    good = brute_force_vanka(K, M, A, B, cells, x)
    bad = brute_force_vanka(K, M, A.T, B.T, cells, x)  # transposed time coupling
    got = P.level_apply(l, 1, x)
    assert np.linalg.norm(got - good) < 1e-12 * np.linalg.norm(good)
    assert np.linalg.norm(got - bad) > 1e-3 * np.linalg.norm(good)


# ------------------------------------------------------------- h-transfer
def gll_points(p):
    """Gauss-Lobatto points on [0, 1]: the end points and the roots of P_p'."""
    inner = np.polynomial.legendre.Legendre.basis(p).deriv().roots() if p > 1 else np.array([])
    return np.concatenate([[0.0], np.sort((inner + 1.0) / 2.0), [1.0]])


def logical_coords(info, p, dim):
    """Logical (unmapped) coordinates of the lexicographic nodes of a level:
    the transfer acts on the reference cells and does not see the geometry."""
    nc = [info["nc0"], info["nc1"], info["nc2"]][:dim]
    g = gll_points(p)
    axes = []
    for d in range(dim):
        a = np.concatenate([(c + g[:-1]) / nc[d] for c in range(nc[d])] + [[1.0]])
        axes.append(a)
    grids = np.meshgrid(*axes[::-1], indexing="ij")[::-1]
    return np.stack([q.reshape(-1) for q in grids], axis=1)


def q_poly(rng, p, dim, bubble):
    """A random function of Q_p (degree <= p per variable); with bubble, one
    that vanishes on the boundary of the unit box (prod x_d (1 - x_d) times
    Q_{p-2})."""
    deg = p - 2 if bubble else p
    C = rng.standard_normal([deg + 1] * dim)

    def f(x):
        v = np.zeros(len(x))
        for e in np.ndindex(*C.shape):
            t = np.full(len(x), C[e])
            for d in range(dim):
                t = t * x[:, d] ** e[d]
            v += t
        if bubble:
            for d in range(dim):
                v = v * x[:, d] * (1 - x[:, d])
        return v
    return f


SPACE_CASES = [(2, 1, 0), (2, 2, 0), (2, 3, 0), (2, 4, 63), (2, 2, 63), (3, 1, 0), (3, 2, 0),
               (3, 3, 63), (3, 4, 0)]


@pytest.mark.parametrize("dim,p,dirichlet", SPACE_CASES)
def test_space_prolongation_reproduces_qp(dim, p, dirichlet):
    refine = 3 if dim == 2 else 2
    # dG(1) with r_min = 1 and one step per slab: every transfer is spatial
    spec = dict(dim=dim, refine=refine, p=p, type=2, r=1, nts=1, r_min=1, dirichlet=dirichlet)
    P = _problem(spec, perturb=0.2, kappa=True)
    rng = np.random.default_rng(5 + p)
    bubble = dirichlet != 0
    if bubble and p < 2:
        pytest.skip("no nonzero Q_1 function vanishes on the boundary")
    for l in range(1, P.n_levels):
        fi, ci = P.level_info(l), P.level_info(l - 1)
        assert fi["mg_type"] == 3
        nt = fi["nt"]
        xf, xc = logical_coords(fi, p, dim), logical_coords(ci, p, dim)
        fs = [q_poly(rng, p, dim, bubble) for _ in range(nt)]
        uc = np.concatenate([f(xc) for f in fs])
        uf = np.concatenate([f(xf) for f in fs])
        got = P.level_apply(l, 2, uc, dst=np.zeros(fi["n_nodes"] * nt))
        assert np.abs(got - uf).max() < 1e-11 * np.abs(uf).max(), l


@pytest.mark.parametrize("dim,p,dirichlet", [(2, 2, 63), (2, 3, 5), (3, 2, 63), (3, 1, 0)])
def test_space_restriction_is_prolongation_transpose(dim, p, dirichlet):
    spec = dict(dim=dim, refine=2 if dim == 2 else 1, p=p, type=2, r=0, nts=1, r_min=0,
                dirichlet=dirichlet)
    P = _problem(spec, perturb=0.1)
    l = P.n_levels - 1
    fi, ci = P.level_info(l), P.level_info(l - 1)
    nf, ncn = fi["n_nodes"], ci["n_nodes"]
    Pm = np.stack([P.level_apply(l, 2, np.eye(ncn)[j], dst=np.zeros(nf)) for j in range(ncn)], 1)
    Rm = np.stack([P.level_apply(l, 3, np.eye(nf)[j], dst=np.zeros(ncn)) for j in range(nf)], 1)
    assert np.array_equal(Rm, Pm.T)
    # Dirichlet rows and columns carry nothing (homogeneous constraints)
    xf = logical_coords(fi, p, dim)
    xc = logical_coords(ci, p, dim)

    def on_dirichlet(x):
        m = np.zeros(len(x), bool)
        for d in range(dim):
            if dirichlet >> (2 * d) & 1:
                m |= np.abs(x[:, d]) < 1e-14
            if dirichlet >> (2 * d + 1) & 1:
                m |= np.abs(x[:, d] - 1) < 1e-14
        return m
    assert not Pm[on_dirichlet(xf), :].any()
    assert not Pm[:, on_dirichlet(xc)].any()
    # every free fine node receives the embedding weights, which sum to 1
    free = ~on_dirichlet(xf)
    inner_c = ~on_dirichlet(xc)
    if dirichlet == 0:
        assert np.allclose(Pm.sum(1), 1.0, atol=1e-13)
    assert Pm[free][:, inner_c].any()


# ------------------------------------------------------ relaxation estimate
OMEGA_CASES = {
    "2d_q2_dg1": (dict(dim=2, refine=2, p=2, type=2, r=1, nts=2, r_min=0), 0.0, False),
    "2d_q3_dg2": (dict(dim=2, refine=2, p=3, type=2, r=2, nts=1, r_min=0), 0.2, True),
    "2d_q2_cg2": (dict(dim=2, refine=2, p=2, type=1, r=2, nts=2, r_min=1), 0.0, False),
    "3d_q2_dg1": (dict(dim=3, refine=1, p=2, type=2, r=1, nts=1, r_min=0), 0.2, True),
}


def dense_op(P, l, what):
    info = P.level_info(l)
    n = info["n_nodes"] * info["nt"]
    return np.stack([P.level_apply(l, what, np.eye(n)[j]) for j in range(n)], 1)


@pytest.mark.parametrize("name", list(OMEGA_CASES))
def test_relaxation_estimate_against_dense_eigenvalues(name):
    """After 20 steps the power iteration's Rayleigh quotient lies within
    [0.75, 1.1] of the spectral radius of V A (measured 0.80-1.05 on these
    levels: V A is not symmetric, its dominant eigenvalues come in complex
    pairs close to the real axis).  The relaxation parameter is
    2 / (1.9 * 1.15 * estimate)."""
    spec, perturb, kappa = OMEGA_CASES[name]
    P = _problem(spec, perturb, kappa)
    checked = 0
    for l in range(1, P.n_levels):
        info = P.level_info(l)
        n = info["n_nodes"] * info["nt"]
        if n > 700:
            continue
        lam_hat = P.level_lambda(l)
        if lam_hat == 0.0:
            assert P.omega(l) == 1.0
            continue
        VA = dense_op(P, l, 1) @ dense_op(P, l, 0)
        rho = np.abs(np.linalg.eigvals(VA)).max()
        assert 0.75 * rho <= lam_hat <= 1.1 * rho, (l, lam_hat, rho)
        lam = 1.15 * lam_hat
        assert abs(P.omega(l) - 2.0 / (0.9 * lam + lam)) < 1e-14 * P.omega(l)
        checked += 1
    assert checked >= 1
    assert P.omega(0) == 1.0  # gmg.h:795-796


# ------------------------------------------------- composition and Krylov
def test_space_time_operator_is_the_kronecker_sum():
    """SystemMatrix::vmult (operators.h:94-122) on block vectors: dst_j =
    sum_i Alpha(j,i) K src_i + Beta(j,i) M src_i, with K and M the assembled
    spatial matrices (constrained rows and columns read as zero) and the
    time weights of fe_time.h; Tvmult with the transposed weights."""
    spec = dict(dim=2, refine=2, p=3, type=2, r=2, nts=2, r_min=0)
    P = _problem(spec, 0.2, True)
    l = P.n_levels - 1
    K, M = P.level_matrix(l, 0), P.level_matrix(l, 1)
    info = P.level_info(l)
    n = info["n_nodes"]
    free = np.ones(n)
    xs = logical_coords(info, spec["p"], 2)
    on_b = np.any((np.abs(xs) < 1e-14) | (np.abs(xs - 1) < 1e-14), axis=1)
    free[on_b] = 0.0
    D = np.diag(free)
    Kf, Mf = D @ K @ D, D @ M @ D
    A, B, _, _ = orc.fe_time_weights(orc.DG, 2, P.tau, 2)
    x = np.random.default_rng(2).standard_normal(n * A.shape[0])
    ref = (np.kron(A, Kf) + np.kron(B, Mf)) @ x
    assert np.linalg.norm(P.fine_vmult(x) - ref) < 1e-13 * np.linalg.norm(ref)
    reft = (np.kron(A.T, Kf) + np.kron(B.T, Mf)) @ x
    assert np.linalg.norm(P.fine_vmult(x, transpose=True) - reft) < 1e-13 * np.linalg.norm(reft)


def test_time_transfers_are_kronecker_products():
    """tensorproduct_add (operators.h:19-69) with the fe_time.h matrices:
    prolongation dst += (P (x) I) src, restriction dst += (P^T (x) I) src,
    for tau (P = time_prolongation) and k levels (P = time_projection)."""
    spec = dict(dim=2, refine=1, p=2, type=2, r=2, nts=2, r_min=0)
    P = _problem(spec)
    rng = np.random.default_rng(12)
    seen = set()
    for l in range(1, P.n_levels):
        fi, ci = P.level_info(l), P.level_info(l - 1)
        kind = fi["mg_type"]
        if kind == 3:
            continue
        Pt = (orc.time_prolongation(orc.DG, fi["r"], fi["nts"]) if kind == 1 else
              orc.time_projection(orc.DG, fi["r"] - 1, fi["r"], fi["nts"]))
        n = fi["n_nodes"]
        I = np.eye(n)
        xc = rng.standard_normal(ci["n_nodes"] * ci["nt"])
        y0 = rng.standard_normal(n * fi["nt"])
        got = P.level_apply(l, 2, xc, dst=y0.copy())
        assert np.allclose(got, y0 + np.kron(Pt, I) @ xc, rtol=0, atol=1e-13)
        xf = rng.standard_normal(n * fi["nt"])
        c0 = rng.standard_normal(ci["n_nodes"] * ci["nt"])
        got = P.level_apply(l, 3, xf, dst=c0.copy())
        assert np.allclose(got, c0 + np.kron(Pt.T, I) @ xf, rtol=0, atol=1e-13)
        seen.add(kind)
    assert seen == {1, 2}


def test_fgmres_residuals_are_minimal_over_the_krylov_space():
    """SolverFGMRES (time_integrators.h:174) with a fixed (linear)
    preconditioner P is right-preconditioned GMRES: its j-th residual is
    min_y |r0 - A P V_j y| over the Krylov space V_j = span{r0, (AP) r0, ...},
    computed here by dense least squares, independently of the Arnoldi
    process and of how the solver updates its residual estimate."""
    spec = dict(dim=2, refine=1, p=2, type=2, r=1, nts=1, r_min=0, reltol=1e-10,
                abstol=1e-300)
    P = _problem(spec)
    info = P.level_info(P.n_levels - 1)
    n = info["n_nodes"] * info["nt"]
    A = np.stack([P.fine_vmult(np.eye(n)[j]) for j in range(n)], 1)
    Pm = np.stack([P.gmg_vmult(np.eye(n)[j]) for j in range(n)], 1)
    xs = logical_coords(info, 2, 2)
    on_b = np.any((np.abs(xs) < 1e-14) | (np.abs(xs - 1) < 1e-14), axis=1)
    b = A @ np.random.default_rng(5).standard_normal(n)
    b[np.tile(on_b, info["nt"])] = 0.0
    rc, x, its, hist = P.fgmres(b, np.zeros(n))
    assert rc == 0 and its >= 2
    AP = A @ Pm
    r0 = b.copy()
    K = [r0 / np.linalg.norm(r0)]
    for j in range(1, its + 1):
        # residual after j steps: best combination of AP K_j
        Q, _ = np.linalg.qr(np.stack(K, 1))
        W = AP @ Q
        y, *_ = np.linalg.lstsq(W, r0, rcond=None)
        rj = np.linalg.norm(r0 - W @ y)
        assert abs(rj - hist[j]) <= 1e-8 * hist[0] + 1e-6 * rj, (j, rj, hist[j])
        K.append(AP @ K[-1])
        K[-1] /= np.linalg.norm(K[-1])
    assert np.linalg.norm(b - A @ x) <= 1.01e-10 * np.linalg.norm(b)
