"""The Gaussian-bump source (spec source=3 of the library, rhs=3 of the
oracle) pinned by a restatement of its generator and by closed-form
integrals.

The benchmark workloads (3D heat, 32^3 Q3/dG(2) and 64^3 Q4/dG(3)) force the
heat equation with seeded Gaussian bumps: bump b has centre c_b (uniform in
[0.2, 0.8]^3) and amplitude a_b (uniform in [0.5, 1.5]) from the 64-bit
Mersenne twister seeded with `seed`, each variate being the top 53 bits of
one output divided by 2^53; h(x) = sum_b a_b exp(-|x - c_b|^2 / (2 0.1^2)),
with unit time profile.  It enters the slab right-hand side through
assemble_force (tests/include/time_integrators.h:60-87 of
nlsmrg/dealii-stfem) as the load vector F_i = int h phi_i.

* The generator: mt19937_64 restated here from its published definition
  (Matsumoto & Nishimura; the parameters of the C++ standard's
  mt19937_64), checked against the standard's known answer (the 10000th
  output of the default-seeded engine is 9981545732273789042).
* The load vector: without Dirichlet faces the FE basis sums to one and
  reproduces x, so sum_i F_i = int h and sum_i x_i F_i = int x h, both in
  closed form (erf); Gauss(p+1) quadrature of the bumps on a 32 x 32 Q2 mesh
  meets them to 1e-9.
"""
import math

import numpy as np
import pytest

import oracle as orc

MASK = (1 << 64) - 1


class MT19937_64:
    """64-bit Mersenne twister (w, n, m, r) = (64, 312, 156, 31)."""
    N, M = 312, 156
    A = 0xB5026F5AA96619E9
    UPPER, LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed=5489):
        self.mt = [0] * self.N
        self.mt[0] = seed & MASK
        for i in range(1, self.N):
            prev = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & MASK
        self.i = self.N

    def _twist(self):
        mt = self.mt
        for k in range(self.N):
            y = (mt[k] & self.UPPER) | (mt[(k + 1) % self.N] & self.LOWER)
            v = mt[(k + self.M) % self.N] ^ (y >> 1)
            if y & 1:
                v ^= self.A
            mt[k] = v
        self.i = 0

    def __call__(self):
        if self.i >= self.N:
            self._twist()
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & MASK


def test_mt19937_64_known_answer():
    g = MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def bumps(n, seed):
    g = MT19937_64(seed)
    out = []
    for _ in range(n):
        u = [(g() >> 11) * 2.0 ** -53 for _ in range(4)]
        out.append((0.2 + 0.6 * u[0], 0.2 + 0.6 * u[1], 0.2 + 0.6 * u[2], 0.5 + u[3]))
    return out


def integrals(bs, dim, sigma=0.1):
    """int h and int x_d h over the unit box, per coordinate d."""
    s2 = sigma * math.sqrt(2.0)

    def i0(c):
        return sigma * math.sqrt(math.pi / 2) * (math.erf((1 - c) / s2) + math.erf(c / s2))

    def i1(c):
        return c * i0(c) + sigma ** 2 * (math.exp(-c * c / (2 * sigma ** 2)) -
                                          math.exp(-(1 - c) ** 2 / (2 * sigma ** 2)))
    tot, mom = 0.0, np.zeros(dim)
    for b in bs:
        f0 = [i0(b[d]) for d in range(dim)]
        tot += b[3] * np.prod(f0)
        for d in range(dim):
            mom[d] += b[3] * i1(b[d]) * np.prod([f0[e] for e in range(dim) if e != d])
    return tot, mom


def node_coords_uniform(dim, n_cells, p):
    gll, _ = orc.quadrature(1, p + 1)
    ax = np.concatenate([(c + gll[:-1]) / n_cells for c in range(n_cells)] + [[1.0]])
    grids = np.meshgrid(*([ax] * dim)[::-1], indexing="ij")[::-1]
    return np.stack([g.reshape(-1) for g in grids], axis=1)


@pytest.mark.parametrize("dim,refine,p,n_bumps,seed", [(2, 5, 2, 4, 7), (2, 4, 3, 3, 11),
                                                       (3, 3, 2, 4, 7)])
def test_bump_load_vector_moments(dim, refine, p, n_bumps, seed):
    P = orc.Problem(dict(dim=dim, refine=refine, p=p, type=2, r=0, nts=1, rhs=3,
                         n_bumps=n_bumps, seed=seed, dirichlet=0, mg_space=0,
                         space_time_mg=0))
    F = P.load_vector(0.0)
    tot, mom = integrals(bumps(n_bumps, seed), dim)
    x = node_coords_uniform(dim, 1 << refine, p)
    tol = 1e-9 if dim == 2 else 1e-6
    assert abs(F.sum() - tot) < tol * tot, (F.sum(), tot)
    for d in range(dim):
        assert abs(x[:, d] @ F - mom[d]) < tol * tot, (d, x[:, d] @ F, mom[d])
    # time profile 1: the same vector at any time
    assert np.array_equal(P.load_vector(0.37), F)
