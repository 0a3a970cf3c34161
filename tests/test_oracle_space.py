"""The oracle's spatial operators pinned by closed forms and invariants that do
not depend on how the oracle evaluates them.

MatrixFreeOperator (tests/include/operators.h:487-548 of nlsmrg/dealii-stfem)
integrates FE_Q(p) functions on Q1-mapped cells with Gauss(p+1) quadrature.
On any mesh of the unit box (perturbed or not) and without constraints:

  * sum_ij M_ij = |domain| = 1 (the basis sums to one);
  * K 1 = 0 (gradients of a constant vanish), K and M symmetric;
  * u^T M 1 = int u for u = x*y (2D) / x*y*z (3D): mapped Q_p with p >= 2
    holds these exactly and Gauss(p+1) integrates u * det J exactly;
  * u^T K u = sum_c kappa_c |c| for u = x (exact on every mesh since the
    Q1 map reproduces affine functions), which pins the per-cell
    coefficient and the metric J^{-T}; |c| from the shoelace formula;
  * Q1 on one Cartesian cell of size h: the 1D mass h/6 [2 1; 1 2]
    tensorised.

The averaged coarse-level coefficients (spec key coarse_coef) reduce to the
sampled one when the finest-level coefficient is constant on every coarse
cell; that is checked on the level operators.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import box_vertices, log_uniform


def csr_dense(rp, ci, v):
    n = len(rp) - 1
    A = np.zeros((n, n))
    for i in range(n):
        A[i, ci[rp[i]:rp[i + 1]]] += v[rp[i]:rp[i + 1]]
    return A


def fine_KM(spec, X=None, kappa=None, sparse=False):
    P = orc.Problem(dict(spec, dirichlet=0), coords=X, kappa=kappa)
    if sparse:
        import scipy.sparse as sp
        mats = []
        for w in (0, 1):
            rp, ci, v = P.fine_matrix(w)
            mats.append(sp.csr_matrix((v, ci, rp), shape=(len(rp) - 1, len(rp) - 1)))
        return P, mats[0], mats[1]
    K = csr_dense(*P.fine_matrix(0))
    M = csr_dense(*P.fine_matrix(1))
    return P, K, M


def node_coords(dim, nc, p, X):
    """Support points of FE_Q(p): GLL points mapped by each cell's Q1 map,
    lexicographic node numbering (x fastest)."""
    gll, _ = orc.quadrature(1, p + 1)
    nn = [nc[d] * p + 1 for d in range(dim)]
    nv = [nc[d] + 1 for d in range(dim)]
    out = np.zeros((int(np.prod(nn)), dim))
    for cell in np.ndindex(*nc[::-1]):
        c = cell[::-1]
        V = []
        for v in range(1 << dim):
            idx = [c[d] + ((v >> d) & 1) for d in range(dim)]
            vid = idx[0] + nv[0] * (idx[1] + (nv[1] * idx[2] if dim == 3 else 0))
            V.append(X[vid])
        for loc in np.ndindex(*([p + 1] * dim)):
            a = loc[::-1]
            xi = [gll[a[d]] for d in range(dim)]
            x = np.zeros(dim)
            for v in range(1 << dim):
                w = 1.0
                for d in range(dim):
                    w *= xi[d] if (v >> d) & 1 else 1 - xi[d]
                x += w * V[v]
            g = [c[d] * p + a[d] for d in range(dim)]
            out[g[0] + nn[0] * (g[1] + (nn[1] * g[2] if dim == 3 else 0))] = x
    return out


def cell_volumes(dim, nc, X):
    nv = [nc[d] + 1 for d in range(dim)]
    vol = []
    for cell in np.ndindex(*nc[::-1]):
        c = cell[::-1]
        if dim == 2:
            def vx(i, j):
                return X[(c[0] + i) + nv[0] * (c[1] + j)]
            q = [vx(0, 0), vx(1, 0), vx(1, 1), vx(0, 1)]
            vol.append(0.5 * abs(sum(q[k][0] * q[(k + 1) % 4][1] - q[(k + 1) % 4][0] * q[k][1]
                                     for k in range(4))))
        else:
            # trilinear hexahedron: det J has degree <= 2 per reference
            # variable, so 2-point Gauss per direction integrates it exactly
            g = 0.5 + np.array([-1.0, 1.0]) / (2.0 * np.sqrt(3.0))
            V = np.array([X[(c[0] + (v & 1)) + nv[0] * ((c[1] + ((v >> 1) & 1)) +
                                                        nv[1] * (c[2] + ((v >> 2) & 1)))]
                          for v in range(8)])
            tot = 0.0
            for xi in np.array(np.meshgrid(g, g, g)).reshape(3, -1).T:
                J = np.zeros((3, 3))
                for v in range(8):
                    b = [(v >> e) & 1 for e in range(3)]
                    for e in range(3):
                        gr = 1.0
                        for f in range(3):
                            if f == e:
                                gr *= 1.0 if b[f] else -1.0
                            else:
                                gr *= xi[f] if b[f] else 1 - xi[f]
                        J[:, e] += gr * V[v]
                tot += 0.125 * abs(np.linalg.det(J))
            vol.append(tot)
    return np.array(vol)


CASES = [(2, 1, 0.0), (2, 2, 0.25), (2, 3, 0.25), (2, 4, 0.2), (3, 1, 0.0), (3, 2, 0.2),
         (3, 3, 0.2), (3, 4, 0.15), (3, 5, 0.1)]


@pytest.mark.parametrize("dim,p,perturb", CASES)
def test_mass_and_stiffness_invariants(dim, p, perturb):
    refine = 2 if dim == 2 else 1
    nc = [1 << refine] * dim
    X = box_vertices(dim, nc, perturb=perturb, seed=3) if perturb else None
    spec = dict(dim=dim, refine=refine, p=p, type=2, r=0, nts=1)
    _, K, M = fine_KM(spec, X)
    one = np.ones(len(M))
    assert abs(one @ M @ one - 1.0) < 1e-13
    assert np.abs(K @ one).max() < 1e-12 * np.abs(K).max()
    assert np.allclose(K, K.T, atol=1e-14 * np.abs(K).max())
    assert np.allclose(M, M.T, atol=1e-15)
    if p >= 2:
        # x*y (2D; degree 2 per reference variable on Q1-mapped cells) and
        # x*y*z (3D, degree 3: in the space from p = 3 on) are interpolated
        # exactly and integrated exactly by Gauss(p+1)
        Xv = X if X is not None else box_vertices(dim, nc, 0.0)
        xs = node_coords(dim, nc, p, Xv)
        k = dim if (dim == 2 or p >= 3) else 2
        u = np.prod(xs[:, :k], axis=1)
        assert abs(u @ M @ one - 0.5 ** k) < 1e-13


@pytest.mark.parametrize("dim,p,perturb", [(2, 1, 0.0), (2, 2, 0.25), (2, 3, 0.2), (3, 1, 0.2),
                                            (3, 3, 0.2), (3, 4, 0.15), (3, 5, 0.1)])
def test_stiffness_energy_with_cell_coefficient(dim, p, perturb):
    refine = 2
    nc = [1 << refine] * dim
    X = box_vertices(dim, nc, perturb=perturb, seed=5)
    kap = log_uniform(int(np.prod(nc)), 1e-2, 1e2, seed=9)
    _, K, _ = fine_KM(dict(dim=dim, refine=refine, p=p, type=2, r=0, nts=1), X, kap, sparse=True)
    u = node_coords(dim, nc, p, X)[:, 0]
    ref = float(np.sum(kap * cell_volumes(dim, nc, X)))
    assert abs(u @ K @ u - ref) < 1e-12 * ref


def test_q1_mass_closed_form():
    # one Cartesian cell [0,1]^2 (subdiv 1, no refinement): M = (1/6 [2 1; 1 2])^{(x)2}
    _, _, M = fine_KM(dict(dim=2, refine=0, p=1, type=2, r=0, nts=1))
    m1 = np.array([[2.0, 1.0], [1.0, 2.0]]) / 6.0
    assert np.allclose(M, np.kron(m1, m1), atol=1e-15)


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_coarse_coefficient_modes_agree_for_coarse_constant_field(mode):
    """A finest-level coefficient that is constant on every level-0 cell is
    represented exactly by every coarse_coef mode, so all level operators
    coincide with the sampled ones (subdivided 2 x 2 box, 2 refinements)."""
    dim, refine, p = 2, 2, 2
    coarse = log_uniform(4, 1e-2, 1e2, seed=2)
    f = 1 << refine
    kap = np.array([coarse[(i // f) + 2 * (j // f)] for j in range(2 * f) for i in range(2 * f)])
    spec = dict(dim=dim, subdiv="2,2", refine=refine, p=p, type=2, r=0, nts=1, mg_float=0)
    a = orc.Problem(dict(spec, coarse_coef=0), kappa=kap)
    b = orc.Problem(dict(spec, coarse_coef=mode), kappa=kap)
    rng = np.random.default_rng(4)
    for l in range(a.n_levels):
        n = a.level_info(l)["n_nodes"] * a.level_info(l)["nt"]
        x = rng.standard_normal(n)
        ya, yb = a.level_apply(l, 0, x), b.level_apply(l, 0, x)
        assert np.allclose(ya, yb, rtol=1e-13, atol=1e-13 * np.abs(ya).max())


def test_coarse_coefficient_means_on_a_checkerboard():
    """Two values lo, hi in a checkerboard of finest cells (Q2, Cartesian):
    on the next coarser level every coarse cell holds two of each.  With
    Gauss(3) points the shares are [0, m0), [m0, m1), [m1, 1] per direction
    (m = midpoints), so the corner points see one finest cell and the others
    see lo and hi with equal overlap.  The coarse stiffness energy of u = x
    is then sum_q w_q kappa_q with kappa_q the chosen mean of its cells:
    corners 2 (lo + hi) w0^2 in total, the rest (1 - 4 w0^2) * mean(lo, hi)."""
    dim, refine, p = 2, 2, 2
    nc = [1 << refine] * dim
    lo, hi = 0.01, 100.0
    kap = np.array([lo if (i + j) % 2 == 0 else hi for j in range(nc[1]) for i in range(nc[0])])
    spec = dict(dim=dim, refine=refine, p=p, type=2, r=0, nts=1, mg_float=0, dirichlet=0)
    w0 = 5.0 / 18.0
    means = {1: np.sqrt(lo * hi), 2: 2.0 / (1.0 / lo + 1.0 / hi), 3: 0.5 * (lo + hi)}
    for mode, mean in means.items():
        P = orc.Problem(dict(spec, coarse_coef=mode), kappa=kap)
        P1 = orc.Problem(dict(spec, coarse_coef=mode), kappa=np.ones_like(kap))
        l = P.n_levels - 2
        n = P.level_info(l)["n_nodes"]
        nn0 = int(round(np.sqrt(n)))
        x = np.tile(np.arange(nn0) / (nn0 - 1), nn0)  # u = x at the nodes
        A, _ = P.level_time_matrices(l)
        # level operator = Alpha K + Beta M (nt = 1): the M parts cancel
        energy = (x @ P.level_apply(l, 0, x) - x @ P1.level_apply(l, 0, x)) / A[0, 0] + 1.0
        expect = 2.0 * (lo + hi) * w0 * w0 + (1.0 - 4.0 * w0 * w0) * mean
        assert abs(energy - expect) < 1e-11 * hi, (mode, energy, expect)


def test_levels_without_free_nodes():
    """Q1 on a mesh whose coarsest level is one cell with Dirichlet faces: that
    level has no free node, its relaxation parameter is 1 (nothing to
    estimate) and the slab solve still converges."""
    spec = dict(dim=3, refine=2, type=2, r=0, nts=2, p=1, r_min=0, vanka_factored=1)
    P = orc.Problem(spec)
    om = [P.omega(l) for l in range(P.n_levels)]
    assert all(np.isfinite(om)), om
    n, nt = P.level_info(P.n_levels - 1)["n_nodes"], P.level_info(P.n_levels - 1)["nt"]
    rc, x, its = orc.solve_slab(P, np.zeros(n * nt), np.zeros(n), 0.0)
    assert rc == 0 and np.all(np.isfinite(x)), (rc, its)
