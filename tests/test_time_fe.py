"""Time finite element algebra of the HIP library (host side, via the C ABI)
against the oracle and closed forms.  The oracle itself is pinned to the
reference's tp_02 / transfer_02 transcripts in test_oracle_goldens.py."""
import numpy as np
import pytest

import oracle as orc
import stfem_amd as st


@pytest.mark.parametrize("kind,n", [(k, n) for k in range(3) for n in range(1, 8)
                                    if not (k == 1 and n < 2)])
def test_quadrature_matches_oracle(kind, n):
    x, w = st.quadrature(kind, n)
    xo, wo = orc.quadrature(kind, n)
    assert np.allclose(x, xo, atol=1e-14) and np.allclose(w, wo, atol=1e-14)
    # exactness: Gauss 2n-1, Lobatto 2n-3, Radau 2n-2
    deg = {0: 2 * n - 1, 1: 2 * n - 3, 2: 2 * n - 2}[kind]
    for k in range(deg + 1):
        assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-13


@pytest.mark.parametrize("type_,r,tau,nts", [(t, r, tau, nts) for t in (1, 2) for r in range(0, 5)
                                             for tau in (1.0, 0.125) for nts in (1, 2, 4)
                                             if not (t == 1 and r == 0)])
def test_fe_time_weights_match_oracle(type_, r, tau, nts):
    a = st.fe_time_weights(type_, r, tau, nts)
    b = orc.fe_time_weights(type_, r, tau, nts)
    for x, y in zip(a, b):
        assert np.allclose(x, y, atol=1e-13, rtol=1e-13)


@pytest.mark.parametrize("type_,r,nts", [(t, r, nts) for t in (1, 2) for r in range(0, 5)
                                         for nts in (1, 2, 3, 4) if not (t == 1 and r == 0)])
def test_wave_weights_match_oracle(type_, r, nts):
    a = st.fe_time_weights_wave(type_, r, 1.0, nts)
    b = orc.wave_weights(type_, r, nts)
    for x, y in zip(a, b):
        assert np.allclose(x, y, atol=1e-11, rtol=1e-11)


@pytest.mark.parametrize("type_,r,nts", [(t, r, nts) for t in (1, 2) for r in range(0, 6)
                                         for nts in (2, 4) if not (t == 1 and r == 0)])
def test_transfers_match_oracle(type_, r, nts):
    assert np.allclose(st.time_prolongation(type_, r, nts), orc.time_prolongation(type_, r, nts),
                       atol=1e-13)
    assert np.allclose(st.time_restriction(type_, r, nts), orc.time_restriction(type_, r, nts),
                       atol=1e-12)


@pytest.mark.parametrize("type_,rs,rd,nts", [(t, a, b, n) for t in (1, 2) for a in range(0, 5)
                                             for b in (a - 1, a + 1) for n in (1, 2)
                                             if b >= (1 if t == 1 else 0) and a >= (1 if t == 1 else 0)])
def test_projection_matches_oracle(type_, rs, rd, nts):
    assert np.allclose(st.time_projection(type_, rs, rd, nts), orc.time_projection(type_, rs, rd, nts),
                       atol=1e-12)


def test_dg_closed_forms():
    # Radau(r+1) is exact for the mass matrix: Alpha = tau diag(Radau weights)
    for r in range(0, 6):
        A, B, G, Z = st.fe_time_weights(2, r, 0.5, 1)
        _, w = st.quadrature(2, r + 1)
        assert np.allclose(A, 0.5 * np.diag(w), atol=1e-14)
        x, _ = st.quadrature(2, r + 1)
        # jump vector = Lagrange basis at t=0
        for i in range(r + 1):
            li = np.prod([(0 - x[k]) / (x[i] - x[k]) for k in range(r + 1) if k != i])
            assert abs(G[i, 0] - li) < 1e-12
        # Beta 1 = jump vector (derivative of constants vanishes, jump term stays)
        assert np.allclose(B.sum(axis=1), G[:, 0], atol=1e-12)
    # DG(0): P = [1;1], R = [1/2 1/2]
    assert np.allclose(st.time_prolongation(2, 0, 2), [[1.0], [1.0]])
    assert np.allclose(st.time_restriction(2, 0, 2), [[0.5, 0.5]])


def test_prolongation_reproduces_polynomials():
    # tau prolongation is the exact embedding of the coarse step into two steps
    for type_ in (1, 2):
        for r in range(1, 5):
            P = st.time_prolongation(type_, r, 2)
            xc, _ = st.quadrature(1 if type_ == 1 else 2, r + 1)
            xf = np.concatenate([0.5 * xc, 0.5 + 0.5 * xc])
            if type_ == 1:  # drop the initial value nodes
                xc = xc[1:]
                xf = np.concatenate([0.5 * st.quadrature(1, r + 1)[0][1:],
                                     0.5 + 0.5 * st.quadrature(1, r + 1)[0][1:]])
            for k in range(r + 1):
                f = lambda t: t ** k
                if type_ == 1 and k > 0:
                    # the dropped initial value is zero for t^k, k >= 1
                    assert np.allclose(P @ f(xc), f(xf), atol=1e-12)
                if type_ == 2:
                    assert np.allclose(P @ f(xc), f(xf), atol=1e-12)


def test_mg_sequence_matches_oracle():
    for args in [(3, 1, 0, 2, 1, 0), (5, 3, 0, 4, 1, 0), (1, 4, 0, 4, 1, 0), (4, 2, 1, 2, 1, 1)]:
        assert st.time_mg_sequence(*args) == orc.mg_sequence(*args)
