// stfem oracle: CPU FP64 restatement of the space-time multigrid slab solve.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
// path; it is built into oracle/build/liboracle.so and may only be loaded by
// tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  The
// product library never links or calls it.
//
// It restates, with plain loops and assembled sparse matrices (no sum
// factorization on the parity path), the algorithm of nlsmrg/dealii-stfem:
//
//   time FE algebra ............ tests/include/fe_time.h:43-716
//   space-time operator ........ tests/include/operators.h:19-256 (SystemMatrix)
//   spatial mass / stiffness ... tests/include/operators.h:355-562 (Gauss(p+1),
//                                Q1 mapping, FE_Q(p) on Gauss-Lobatto nodes)
//   Vanka cell-patch smoother .. tests/include/gmg.h:402-540
//   time transfers ............. tests/include/gmg.h:77-136 (R = P^T)
//   space transfers ............ tests/include/gmg.h:35-75, 317-400 (deal.II
//                                MGTwoLevelTransfer: FE embedding, zero Dirichlet)
//   V-cycle + relaxation ....... tests/include/gmg.h:652-859
//   FGMRES / ReductionControl .. tests/include/time_integrators.h:21-180
//   slab loop, rhs, errors ..... tests/tp_01.cc:56-706, tests/transfer_01.cc:364-744,
//                                tests/include/exact_solution.h:18-314
//
// deal.II-internal semantics restated here (deal.II is a third-party
// dependency that is not in this tree; behaviour per deal.II 9.5):
//   * MatrixFree with homogeneous constraints: constrained dofs read as zero,
//     not written, so the operator is [A_ff 0; 0 0].
//   * MatrixFreeTools::compute_matrix + AffineConstraints::distribute_local_
//     to_global: constrained rows/columns eliminated, diagonal entry of a
//     constrained dof = sum over its cells of |A_cell(i,i)|.
//   * PreconditionRelaxation step_operations, MGSmootherPrecondition (variable
//     steps 2^(L-l)), MGCoarseGridApplySmoother, Multigrid::level_v_step with
//     prolongate_and_add into the level solution.
//   * PreconditionChebyshev power iteration (Rayleigh quotient, 20 its),
//     omega = 2/(alpha + lambda_max), alpha = 0.9 lambda_max.  The start
//     vector and any safety factor are deal.II-internal; the defaults here
//     (constant start vector, factor 1.15 on the estimate) are the best fit
//     to the golden iteration counts (other variants selectable by spec).
//     This part is parity-unpinned beyond those counts.
//   * SolverFGMRES (max basis 30, modified Gram-Schmidt, least squares on the
//     j x (j-1) Hessenberg block, iteration counted per projected solve).
// These are pinned only through the golden iteration counts and error norms
// of tests/transfer_01.output and tests/tp_01.output.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc
{
  enum TimeType
  {
    CGP = 1,
    DG  = 2
  };
  enum MGType
  {
    MG_TAU  = 1,
    MG_K    = 2,
    MG_NONE = 3
  };

  // ------------------------------------------------------------------ dense
  struct Mat
  {
    int                 m = 0, n = 0;
    std::vector<double> a;
    Mat() = default;
    Mat(int m_, int n_)
      : m(m_)
      , n(n_)
      , a(size_t(m_) * n_, 0.0)
    {}
    double &operator()(int i, int j) { return a[size_t(i) * n + j]; }
    double  operator()(int i, int j) const { return a[size_t(i) * n + j]; }
    bool    all_zero() const
    {
      for (double v : a)
        if (v != 0.0)
          return false;
      return true;
    }
  };

  static Mat transpose(const Mat &A)
  {
    Mat T(A.n, A.m);
    for (int i = 0; i < A.m; ++i)
      for (int j = 0; j < A.n; ++j)
        T(j, i) = A(i, j);
    return T;
  }

  static Mat matmul(const Mat &A, const Mat &B)
  {
    Mat C(A.m, B.n);
    for (int i = 0; i < A.m; ++i)
      for (int k = 0; k < A.n; ++k)
        for (int j = 0; j < B.n; ++j)
          C(i, j) += A(i, k) * B(k, j);
    return C;
  }

  // Gauss-Jordan inverse with partial pivoting, in precision T.
  template <typename T>
  static bool invert_inplace(std::vector<T> &A, int n)
  {
    std::vector<int> piv(n);
    std::iota(piv.begin(), piv.end(), 0);
    std::vector<T> inv(size_t(n) * n, T(0));
    for (int i = 0; i < n; ++i)
      inv[size_t(i) * n + i] = T(1);
    for (int c = 0; c < n; ++c)
      {
        int r = c;
        T   best = std::abs(A[size_t(c) * n + c]);
        for (int k = c + 1; k < n; ++k)
          if (std::abs(A[size_t(k) * n + c]) > best)
            best = std::abs(A[size_t(k) * n + c]), r = k;
        if (best == T(0))
          return false;
        if (r != c)
          for (int j = 0; j < n; ++j)
            {
              std::swap(A[size_t(r) * n + j], A[size_t(c) * n + j]);
              std::swap(inv[size_t(r) * n + j], inv[size_t(c) * n + j]);
            }
        const T d = T(1) / A[size_t(c) * n + c];
        for (int j = 0; j < n; ++j)
          {
            A[size_t(c) * n + j] *= d;
            inv[size_t(c) * n + j] *= d;
          }
        for (int k = 0; k < n; ++k)
          if (k != c)
            {
              const T f = A[size_t(k) * n + c];
              if (f != T(0))
                for (int j = 0; j < n; ++j)
                  {
                    A[size_t(k) * n + j] -= f * A[size_t(c) * n + j];
                    inv[size_t(k) * n + j] -= f * inv[size_t(c) * n + j];
                  }
            }
      }
    A.swap(inv);
    return true;
  }

  static Mat inverse(const Mat &A)
  {
    std::vector<double> v = A.a;
    if (!invert_inplace(v, A.m))
      {
        fprintf(stderr, "oracle: singular matrix\n");
        abort();
      }
    Mat R(A.m, A.n);
    R.a = v;
    return R;
  }

  // ------------------------------------------------ 1D polynomials/quadrature
  // Legendre P_n(x) on [-1,1] by the three-term recurrence.
  static long double legendre(int n, long double x)
  {
    if (n == 0)
      return 1.0L;
    long double p0 = 1.0L, p1 = x;
    for (int k = 2; k <= n; ++k)
      {
        long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0             = p1;
        p1             = p2;
      }
    return p1;
  }
  static long double legendre_deriv(int n, long double x)
  {
    // derivative by finite-free recurrence: P'_n = n (x P_n - P_{n-1}) / (x^2-1)
    // avoided at |x|=1 (never evaluated there).
    if (n == 0)
      return 0.0L;
    return n * (x * legendre(n, x) - legendre(n - 1, x)) / (x * x - 1.0L);
  }

  // All roots of f in the open interval (-1,1), found by sign changes on a
  // fine sampling and bisection (robust for the low degrees used here).
  template <typename F>
  static std::vector<long double> roots_open(F f, int expected)
  {
    std::vector<long double> r;
    const int                S = 20000;
    long double              xa = -1.0L + 1e-12L, fa = f(xa);
    for (int s = 1; s <= S; ++s)
      {
        long double xb = -1.0L + 2.0L * (s - 0.318309886L) / S;
        if (s == S)
          xb = 1.0L - 1e-12L;
        long double fb = f(xb);
        if ((fa < 0) != (fb < 0))
          {
            long double lo = xa, hi = xb, flo = fa;
            for (int it = 0; it < 200; ++it)
              {
                long double mid = 0.5L * (lo + hi), fm = f(mid);
                if ((fm < 0) == (flo < 0))
                  lo = mid, flo = fm;
                else
                  hi = mid;
              }
            r.push_back(0.5L * (lo + hi));
          }
        xa = xb;
        fa = fb;
      }
    if ((int)r.size() != expected)
      {
        fprintf(stderr, "oracle: root finder found %d of %d roots\n", (int)r.size(), expected);
        abort();
      }
    return r;
  }

  // Gauss points on [0,1] (deal.II QGauss<1>(n)).
  static std::vector<double> gauss_points(int n)
  {
    auto rr = roots_open([n](long double x) { return legendre(n, x); }, n);
    std::vector<double> p;
    for (auto x : rr)
      p.push_back(double(0.5L * (x + 1.0L)));
    return p;
  }
  static std::vector<double> gauss_weights(int n)
  {
    auto rr = roots_open([n](long double x) { return legendre(n, x); }, n);
    std::vector<double> w;
    for (auto x : rr)
      {
        long double d = legendre_deriv(n, x);
        w.push_back(double(1.0L / ((1.0L - x * x) * d * d)));
      }
    return w;
  }
  // Gauss-Lobatto points on [0,1] (deal.II QGaussLobatto<1>(n)), n >= 2.
  static std::vector<double> lobatto_points(int n)
  {
    std::vector<double> p{0.0};
    if (n > 2)
      {
        auto rr = roots_open([n](long double x) { return legendre_deriv(n - 1, x); }, n - 2);
        for (auto x : rr)
          p.push_back(double(0.5L * (x + 1.0L)));
      }
    p.push_back(1.0);
    return p;
  }
  // Gauss-Radau points on [0,1] with the right end point included
  // (deal.II QGaussRadau<1>(n, right)): roots of P_{n-1} - P_n.
  static std::vector<double> radau_right_points(int n)
  {
    std::vector<double> p;
    if (n > 1)
      {
        auto rr = roots_open(
          [n](long double x) { return (legendre(n - 1, x) - legendre(n, x)) / (1.0L - x); },
                             n - 1);
        for (auto x : rr)
          p.push_back(double(0.5L * (x + 1.0L)));
      }
    p.push_back(1.0);
    return p;
  }

  struct Lagrange
  {
    std::vector<double> x;
    double              value(int i, double t) const
    {
      double v = 1.0;
      for (size_t k = 0; k < x.size(); ++k)
        if ((int)k != i)
          v *= (t - x[k]) / (x[i] - x[k]);
      return v;
    }
    double deriv(int i, double t) const
    {
      double s = 0.0;
      for (size_t k = 0; k < x.size(); ++k)
        if ((int)k != i)
          {
            double v = 1.0 / (x[i] - x[k]);
            for (size_t m = 0; m < x.size(); ++m)
              if ((int)m != i && m != k)
                v *= (t - x[m]) / (x[i] - x[m]);
            s += v;
          }
      return s;
    }
    int size() const { return (int)x.size(); }
  };

  // Exact integral over [a,b] of a polynomial of degree <= 39 given as a
  // function, by 20-point Gauss quadrature.
  template <typename F>
  static double integrate(F f, double a = 0.0, double b = 1.0)
  {
    static const std::vector<double> gp = gauss_points(20);
    static const std::vector<double> gw = gauss_weights(20);
    double                           s  = 0.0;
    for (size_t q = 0; q < gp.size(); ++q)
      s += gw[q] * f(a + (b - a) * gp[q]) * (b - a);
    return s;
  }

  // --------------------------------------------------------- time FE algebra
  // fe_time.h:414-425
  static Lagrange time_basis(int type, int r)
  {
    Lagrange L;
    L.x = (type == CGP) ? lobatto_points(r + 1) : radau_right_points(r + 1);
    return L;
  }

  // fe_time.h:514-555 : {lhs (mass), lhs_der (jump + derivative), jump}
  static std::vector<Mat> dg_weights(int r)
  {
    Lagrange     L = time_basis(DG, r);
    const int    n = r + 1;
    Mat          mass(n, n), der(n, n), jump(n, 1);
    const auto   gq = gauss_points(r + 2);
    const auto   gw = gauss_weights(r + 2);
    for (int i = 0; i < n; ++i)
      {
        jump(i, 0) = L.value(i, 0.0);
        for (int j = 0; j < n; ++j)
          for (size_t q = 0; q < gq.size(); ++q)
            mass(i, j) += gw[q] * L.value(i, gq[q]) * L.value(j, gq[q]);
      }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        {
          der(i, j) += L.value(i, 0.0) * L.value(j, 0.0);
          for (size_t q = 0; q < gq.size(); ++q)
            der(i, j) += gw[q] * L.value(i, gq[q]) * L.deriv(j, gq[q]);
        }
    return {mass, der, jump};
  }

  // fe_time.h:473-512 : {matrix (r x r+1), matrix_der (r x r+1)}
  static std::vector<Mat> cg_weights(int r)
  {
    Lagrange trial = time_basis(CGP, r);
    Lagrange test;
    test.x.assign(trial.x.begin() + 1, trial.x.end());
    const auto gq = gauss_points(r + 2);
    const auto gw = gauss_weights(r + 2);
    Mat        m(r, r + 1), d(r, r + 1);
    for (int i = 0; i < r; ++i)
      for (int j = 0; j < r + 1; ++j)
        for (size_t q = 0; q < gq.size(); ++q)
          {
            m(i, j) += gw[q] * test.value(i, gq[q]) * trial.value(j, gq[q]);
            d(i, j) += gw[q] * test.value(i, gq[q]) * trial.deriv(j, gq[q]);
          }
    return {m, d};
  }

  // fe_time.h:428-471
  static std::vector<Mat> split_lhs_rhs_cg(const std::vector<Mat> &w)
  {
    const int m = w[0].m, n = w[0].n;
    Mat       a(m, n - 1), b(m, n - 1), g(m, 1), z(m, 1);
    for (int i = 0; i < m; ++i)
      {
        for (int j = 1; j < n; ++j)
          {
            a(i, j - 1) = w[0](i, j);
            b(i, j - 1) = w[1](i, j);
          }
        g(i, 0) = -w[0](i, 0);
        z(i, 0) = -w[1](i, 0);
      }
    return {a, b, g, z};
  }

  // fe_time.h:292-349 : {Alpha, Beta, Gamma, Zeta}
  static std::vector<Mat> fe_time_weights(int type, int r, double tau, int nts)
  {
    std::vector<Mat> tmp;
    if (type == CGP)
      {
        tmp = split_lhs_rhs_cg(cg_weights(r));
        for (double &v : tmp[2].a)
          v *= tau;
      }
    else
      {
        auto w = dg_weights(r);
        tmp    = {w[0], w[1], Mat(w[2].m, 1), w[2]};
      }
    for (double &v : tmp[0].a)
      v *= tau;
    const int ni = tmp[0].m, nt = ni * nts;
    Mat       A(nt, nt), B(nt, nt), G(nt, 1), Z(nt, 1);
    for (int it = 0; it < nts; ++it)
      for (int i = 0; i < ni; ++i)
        {
          if (it < nts - 1 && i == ni - 1)
            for (int j = 0; j < ni; ++j)
              {
                A(j + (it + 1) * ni, i + it * ni) = -tmp[2](j, 0);
                B(j + (it + 1) * ni, i + it * ni) = -tmp[3](j, 0);
              }
          for (int j = 0; j < ni; ++j)
            {
              A(i + it * ni, j + it * ni) = tmp[0](i, j);
              B(i + it * ni, j + it * ni) = tmp[1](i, j);
            }
        }
    for (int i = 0; i < ni; ++i)
      {
        G(i, 0) = tmp[type == CGP ? 2 : 3](i, 0);
        Z(i, 0) = tmp[type == CGP ? 3 : 2](i, 0);
      }
    return {A, B, G, Z};
  }

  // fe_time.h:98-246 : {lhs_uK, lhs_uM, rhs_uK, rhs_uM, rhs_vM}
  static std::vector<Mat> fe_time_weights_wave(int type, const Mat &Alpha, const Mat &Beta,
                                               const Mat &Gamma, const Mat &Zeta, int nts)
  {
    const int ni = Alpha.m, ntot = ni * nts;
    Mat       Ai   = inverse(Alpha);
    Mat       BAiB = matmul(Beta, matmul(Ai, Beta));
    Mat       BAiG = matmul(Beta, matmul(Ai, Gamma));
    double    gxai = Gamma(ni - 1, 0) / Alpha(ni - 1, ni - 1);
    Mat       GAiG = Gamma;
    for (double &v : GAiG.a)
      v *= gxai;
    Mat Brow(1, Beta.n);
    for (int j = 0; j < Beta.n; ++j)
      Brow(0, j) = Beta(Beta.m - 1, j);
    Mat GAiB = matmul(Gamma, Brow);
    for (double &v : GAiB.a)
      v /= Alpha(ni - 1, ni - 1);
    Mat r0(ntot, ntot), r1(ntot, ntot), r2(ntot, 1), r3(ntot, 1), r4(ntot, 1);
    if (type == CGP)
      {
        Mat BAiZ = matmul(Beta, matmul(Ai, Zeta));
        Mat ZmBAiG = Zeta;
        for (size_t k = 0; k < ZmBAiG.a.size(); ++k)
          ZmBAiG.a[k] -= BAiG.a[k];
        Mat ZmBAiB = matmul(ZmBAiG, Brow);
        for (double &v : ZmBAiB.a)
          v /= Alpha(ni - 1, ni - 1);
        double zxai = Zeta(ni - 1, 0) / Alpha(ni - 1, ni - 1);
        for (int it = 0; it < nts; ++it)
          for (int jt = 0; jt <= it; ++jt)
            for (int i = 0; i < ni; ++i)
              {
                if (it == 0 && jt == 0)
                  {
                    r2(i, 0) = Gamma(i, 0);
                    r3(i, 0) = BAiZ(i, 0);
                    r4(i, 0) = ZmBAiG(i, 0);
                  }
                else if (jt == 0)
                  {
                    r3(i + it * ni, 0) = -zxai * std::pow(gxai, it - 1) * ZmBAiG(i, 0);
                    r4(i + it * ni, 0) = std::pow(gxai, it) * ZmBAiG(i, 0);
                  }
                if (it == jt + 1)
                  {
                    r0(i + it * ni, ni - 1 + jt * ni) = -Gamma(i, 0);
                    r1(i + it * ni, ni - 1 + jt * ni) = -BAiZ(i, 0);
                  }
                if (it == jt)
                  for (int j = 0; j < ni; ++j)
                    {
                      r0(i + it * ni, j + it * ni) = Alpha(i, j);
                      r1(i + it * ni, j + it * ni) = BAiB(i, j);
                    }
                else
                  for (int j = 0; j < ni; ++j)
                    r1(i + it * ni, j + jt * ni) +=
                      -std::pow(gxai, it - jt - 1) * ZmBAiB(i, j) +
                      ((it > 1 && it - 1 > jt && j == ni - 1) ?
                         std::pow(gxai, it - jt - 2) * zxai * ZmBAiG(i, 0) :
                         0.0);
              }
      }
    else
      {
        for (int it = 0; it < nts; ++it)
          for (int i = 0; i < ni; ++i)
            {
              if (it == 0)
                {
                  r3(i, 0) = BAiG(i, 0);
                  r4(i, 0) = Gamma(i, 0);
                }
              if (it == 1)
                r3(ni + i, 0) = -GAiG(i, 0);
              if (it < nts - 1)
                for (int j = 0; j < ni; ++j)
                  r1(j + (it + 1) * ni, i + it * ni) =
                    -GAiB(j, i) - (i == ni - 1 ? BAiG(j, 0) : 0.0);
              if (it < nts - 2 && i == ni - 1)
                for (int j = 0; j < ni; ++j)
                  r1(j + (it + 2) * ni, i + it * ni) = GAiG(j, 0);
              for (int j = 0; j < ni; ++j)
                {
                  r0(i + it * ni, j + it * ni) = Alpha(i, j);
                  r1(i + it * ni, j + it * ni) = BAiB(i, j);
                }
            }
      }
    return {r0, r1, r2, r3, r4};
  }

  // Embedding of the coarse 1D element into child c (interpolation at the
  // child's nodes): E(i,j) = phi_j((x_i + c)/2).  (FE_DGQArbitraryNodes /
  // FE_Q<1> get_prolongation_matrix, nodes in increasing order.)
  static Mat child_embedding(const Lagrange &L, int c)
  {
    const int n = L.size();
    Mat       E(n, n);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        E(i, j) = L.value(j, 0.5 * (L.x[i] + c));
    return E;
  }
  // L2 projection of child c's space onto the coarse space:
  // R = Mc^{-1} * int_{child} phi^coarse phi^child (in parent coordinates).
  static Mat child_l2_restriction(const Lagrange &L, int c)
  {
    const int n = L.size();
    Mat       Mc(n, n), Mx(n, n);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        {
          Mc(i, j) = integrate([&](double t) { return L.value(i, t) * L.value(j, t); });
          Mx(i, j) = integrate(
            [&](double t) { return L.value(i, t) * L.value(j, 2.0 * t - c); }, 0.5 * c,
            0.5 * (c + 1));
        }
    return matmul(inverse(Mc), Mx);
  }
  // Interpolatory (non-additive) restriction of FE_Q<1>: coarse node value
  // taken from the child containing the node.
  static Mat child_interp_restriction(const Lagrange &L, int c)
  {
    const int n = L.size();
    Mat       R(n, n);
    for (int i = 0; i < n; ++i)
      {
        const double x = L.x[i];
        const bool   in = (c == 0) ? (x <= 0.5 + 1e-14) : (x >= 0.5 - 1e-14);
        if (!in)
          continue;
        for (int j = 0; j < n; ++j)
          R(i, j) = L.value(j, 2.0 * x - c);
      }
    return R;
  }

  // fe_time.h:625-669
  static Mat time_prolongation(int type, int r, int nts)
  {
    Lagrange L = time_basis(type, r);
    Mat      E0 = child_embedding(L, 0), E1 = child_embedding(L, 1);
    const int ni = (type == DG) ? r + 1 : r;
    Mat       P(2 * ni, ni);
    if (type == DG)
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < ni; ++j)
          {
            P(i, j)      = E0(i, j);
            P(ni + i, j) = E1(i, j);
          }
    else
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < ni; ++j)
          {
            P(i, j)      = E0(i + 1, j + 1);
            P(ni + i, j) = E1(i + 1, j + 1);
          }
    Mat Pn(ni * nts, ni * nts / 2);
    for (int it = 0; it < nts / 2; ++it)
      for (int i = 0; i < 2 * ni; ++i)
        for (int j = 0; j < ni; ++j)
          Pn(it * 2 * ni + i, it * ni + j) = P(i, j);
    return Pn;
  }

  // fe_time.h:671-716
  static Mat time_restriction(int type, int r, int nts)
  {
    Lagrange L = time_basis(type, r);
    const int ni = (type == DG) ? r + 1 : r;
    Mat       R(ni, 2 * ni);
    if (type == DG)
      {
        Mat R0 = child_l2_restriction(L, 0), R1 = child_l2_restriction(L, 1);
        for (int i = 0; i < ni; ++i)
          for (int j = 0; j < ni; ++j)
            {
              R(i, j)      = R0(i, j);
              R(i, ni + j) = R1(i, j);
            }
      }
    else
      {
        Mat R0 = child_interp_restriction(L, 0), R1 = child_interp_restriction(L, 1);
        for (int i = 0; i < ni; ++i)
          for (int j = 0; j < ni; ++j)
            {
              R(i, j)      = R0(i + 1, j + 1);
              R(i, ni + j) = R1(i + 1, j + 1);
            }
      }
    Mat Rn(ni * nts / 2, ni * nts);
    for (int it = 0; it < nts / 2; ++it)
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < 2 * ni; ++j)
          Rn(it * ni + i, it * 2 * ni + j) = R(i, j);
    return Rn;
  }

  // fe_time.h:567-623 : L2 projection between degree r_src and r_dst,
  // block-diagonal over the steps; for CGP the initial dof row/col dropped.
  static Mat time_projection(int type, int r_src, int r_dst, int nts)
  {
    Lagrange Ls = time_basis(type, r_src), Ld = time_basis(type, r_dst);
    const int ns = r_src + 1, nd = r_dst + 1;
    Mat       Md(nd, nd), Mx(nd, ns);
    for (int i = 0; i < nd; ++i)
      {
        for (int j = 0; j < nd; ++j)
          Md(i, j) = integrate([&](double t) { return Ld.value(i, t) * Ld.value(j, t); });
        for (int j = 0; j < ns; ++j)
          Mx(i, j) = integrate([&](double t) { return Ld.value(i, t) * Ls.value(j, t); });
      }
    Mat       P  = matmul(inverse(Md), Mx);
    const int di = (type == DG) ? r_dst + 1 : r_dst;
    const int si = (type == DG) ? r_src + 1 : r_src;
    const int nD = (type == DG) ? nts * (r_dst + 1) : nts * r_dst + 1;
    const int nS = (type == DG) ? nts * (r_src + 1) : nts * r_src + 1;
    Mat       Pn(nD, nS);
    for (int it = 0; it < nts; ++it)
      for (int i = 0; i < nd; ++i)
        for (int j = 0; j < ns; ++j)
          if (it * di + i < nD && it * si + j < nS)
            Pn(it * di + i, it * si + j) = P(i, j);
    if (type == CGP)
      {
        Mat Q(nD - 1, nS - 1);
        for (int i = 1; i < nD; ++i)
          for (int j = 1; j < nS; ++j)
            Q(i - 1, j - 1) = Pn(i, j);
        return Q;
      }
    return Pn;
  }

  // fe_time.h:43-69
  static std::vector<int> time_mg_sequence(int n_sp_lvl, int k_max, int k_min, int nts,
                                           int nts_min, bool time_before_space)
  {
    int n_k = k_max - k_min;
    int n_t = (int)std::log2((double)(nts / nts_min));
    // lower level type is k
    int              n_ll = n_k, n_ul = n_t;
    std::vector<int> v(n_k + n_t + n_sp_lvl - 1, MG_NONE);
    int              s_lo = time_before_space ? n_sp_lvl - 1 : 0;
    int              s_up = time_before_space ? n_sp_lvl - 1 + n_ll : n_ll;
    std::fill_n(v.begin() + s_lo, n_ll, (int)MG_K);
    std::fill_n(v.begin() + s_up, n_ul, (int)MG_TAU);
    return v;
  }

  // --------------------------------------------------------------- sparse
  struct CSR
  {
    long              n = 0, m = 0;
    std::vector<long> rp;
    std::vector<long> ci;
    std::vector<double> v;
  };

  // ------------------------------------------------------------ mesh / FE
  // Structured box mesh: nc[d] cells per direction, lexicographic vertices
  // (x fastest).  FE_Q(p) dofs on the global Gauss-Lobatto node grid,
  // numbered lexicographically; all box faces carry zero Dirichlet values.
  struct Grid
  {
    int                 dim = 2;
    int                 nc[3] = {1, 1, 1};
    int                 dirichlet = 63; // faces with zero Dirichlet: bit 2d low, 2d+1 high side
    std::vector<double> X;
    long n_cells() const { return (long)nc[0] * nc[1] * (dim == 3 ? nc[2] : 1); }
    long nvx(int d) const { return nc[d] + 1; }
    long vid(int i, int j, int k) const { return i + nvx(0) * (j + nvx(1) * (long)k); }
  };

  struct Shape1D
  {
    int                 p, n;
    std::vector<double> gll, gq, gw;
    std::vector<double> S, D; // S[q*n+a] = phi_a(gq), D[q*n+a] = phi_a'(gq)
    Lagrange            L;
    void init(int p_, int nq)
    {
      p   = p_;
      n   = p + 1;
      gll = lobatto_points(n);
      gq  = gauss_points(nq);
      gw  = gauss_weights(nq);
      L.x = gll;
      S.assign(size_t(nq) * n, 0.0);
      D.assign(size_t(nq) * n, 0.0);
      for (int q = 0; q < nq; ++q)
        for (int a = 0; a < n; ++a)
          {
            S[q * n + a] = L.value(a, gq[q]);
            D[q * n + a] = L.deriv(a, gq[q]);
          }
    }
  };

  struct Space
  {
    Grid    g;
    int     p = 1, n = 2, dim = 2;
    long    N[3] = {1, 1, 1};
    long    ndofs = 0;
    int     nc_local = 0;
    std::vector<char> constrained;
    void init(const Grid &g_, int p_)
    {
      g   = g_;
      p   = p_;
      n   = p + 1;
      dim = g.dim;
      ndofs = 1;
      for (int d = 0; d < 3; ++d)
        {
          N[d] = (d < dim) ? (long)g.nc[d] * p + 1 : 1;
          ndofs *= N[d];
        }
      nc_local = (dim == 3) ? n * n * n : n * n;
      constrained.assign(ndofs, 0);
      for (long k = 0; k < N[2]; ++k)
        for (long j = 0; j < N[1]; ++j)
          for (long i = 0; i < N[0]; ++i)
            {
              const int m = g.dirichlet;
              bool b = (i == 0 && (m & 1)) || (i == N[0] - 1 && (m & 2)) ||
                       (j == 0 && (m & 4)) || (j == N[1] - 1 && (m & 8));
              if (dim == 3)
                b = b || (k == 0 && (m & 16)) || (k == N[2] - 1 && (m & 32));
              constrained[i + N[0] * (j + N[1] * k)] = b;
            }
    }
    void cell_coords(long c, int &cx, int &cy, int &cz) const
    {
      cx = int(c % g.nc[0]);
      cy = int((c / g.nc[0]) % g.nc[1]);
      cz = (dim == 3) ? int(c / ((long)g.nc[0] * g.nc[1])) : 0;
    }
    void cell_dofs(long c, std::vector<long> &idx) const
    {
      int cx, cy, cz;
      cell_coords(c, cx, cy, cz);
      idx.resize(nc_local);
      int l = 0;
      for (int kk = 0; kk < (dim == 3 ? n : 1); ++kk)
        for (int jj = 0; jj < n; ++jj)
          for (int ii = 0; ii < n; ++ii)
            idx[l++] = (cx * p + ii) + N[0] * ((cy * p + jj) + N[1] * (long)(cz * p + kk));
    }
    // Q1 vertex coordinates of cell c, in tensor order (v = a + 2b + 4c).
    void cell_vertices(long c, double V[8][3]) const
    {
      int cx, cy, cz;
      cell_coords(c, cx, cy, cz);
      for (int v = 0; v < (1 << dim); ++v)
        {
          int  a = v & 1, b = (v >> 1) & 1, cc = (v >> 2) & 1;
          long id = g.vid(cx + a, cy + b, cz + cc);
          for (int d = 0; d < dim; ++d)
            V[v][d] = g.X[id * dim + d];
        }
    }
  };

  // Q1 mapping: position, Jacobian J[d][e] = dx_d/dxi_e at reference xi.
  static void q1_map(int dim, const double V[8][3], const double *xi, double *x, double J[3][3])
  {
    for (int d = 0; d < dim; ++d)
      {
        x[d] = 0;
        for (int e = 0; e < dim; ++e)
          J[d][e] = 0;
      }
    for (int v = 0; v < (1 << dim); ++v)
      {
        double N = 1, dN[3];
        for (int e = 0; e < dim; ++e)
          {
            int    bit = (v >> e) & 1;
            double t   = bit ? xi[e] : 1 - xi[e];
            N *= t;
          }
        for (int e = 0; e < dim; ++e)
          {
            double g = 1;
            for (int f = 0; f < dim; ++f)
              {
                int bit = (v >> f) & 1;
                if (f == e)
                  g *= bit ? 1.0 : -1.0;
                else
                  g *= bit ? xi[f] : 1 - xi[f];
              }
            dN[e] = g;
          }
        for (int d = 0; d < dim; ++d)
          {
            x[d] += N * V[v][d];
            for (int e = 0; e < dim; ++e)
              J[d][e] += dN[e] * V[v][d];
          }
      }
  }
  static double det_inv(int dim, double J[3][3], double Ji[3][3])
  {
    if (dim == 2)
      {
        double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        Ji[0][0]   = J[1][1] / det;
        Ji[0][1]   = -J[0][1] / det;
        Ji[1][0]   = -J[1][0] / det;
        Ji[1][1]   = J[0][0] / det;
        return det;
      }
    double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                 J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                 J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    return det;
  }

  // Coefficient: per finest-level cell value (or constant 1 when absent).
  // On a coarser level the coefficient at a quadrature point is, for mode 0,
  // the value of the finest cell containing that point in reference
  // coordinates (a coefficient function sampled at the point, as
  // MatrixFreeOperator::evaluate_coefficient does, operators.h:445-457);
  // modes 1, 2, 3 average over the finest cells that meet the point's share
  // of the coarse cell (per direction the interval between the midpoints of
  // neighbouring Gauss points), weighted by overlap: geometric, harmonic,
  // arithmetic mean.  The averaged modes are an extension of this build for
  // rough coefficients; the reference has only the sampled form.
  struct Coefficient
  {
    int                 dim = 2;
    int                 fine_nc[3] = {1, 1, 1};
    std::vector<double> kappa; // empty: constant 1
    int                 mode = 0;
    std::vector<double> gq; // Gauss points of the quadrature the levels use
    double at(int lev_shift, int cx, int cy, int cz, const double *xi) const
    {
      if (kappa.empty())
        return 1.0;
      const int s = 1 << lev_shift;
      if (mode != 0 && s > 1)
        return averaged(s, cx, cy, cz, xi);
      int       f[3] = {cx, cy, cz};
      for (int d = 0; d < dim; ++d)
        {
          int sub = (int)std::floor(xi[d] * s);
          sub     = std::min(std::max(sub, 0), s - 1);
          f[d]    = f[d] * s + sub;
        }
      long id = f[0] + (long)fine_nc[0] * (f[1] + (long)fine_nc[1] * (dim == 3 ? f[2] : 0));
      return kappa[id];
    }
    double averaged(int s, int cx, int cy, int cz, const double *xi) const
    {
      const int nq = (int)gq.size();
      std::vector<std::pair<int, double>> share[3];
      for (int d = 0; d < 3; ++d)
        {
          if (d >= dim)
            {
              share[d].push_back({0, 1.0});
              continue;
            }
          int q = 0;
          for (int k = 1; k < nq; ++k)
            if (std::abs(gq[k] - xi[d]) < std::abs(gq[q] - xi[d]))
              q = k;
          const double lo = q == 0 ? 0.0 : 0.5 * (gq[q - 1] + gq[q]);
          const double hi = q == nq - 1 ? 1.0 : 0.5 * (gq[q] + gq[q + 1]);
          for (int sub = 0; sub < s; ++sub)
            {
              const double ov = std::min(hi, double(sub + 1) / s) - std::max(lo, double(sub) / s);
              if (ov > 0)
                share[d].push_back({sub, ov});
            }
        }
      double acc = 0.0, wsum = 0.0;
      for (const auto &ez : share[2])
        for (const auto &ey : share[1])
          for (const auto &ex : share[0])
            {
              const long   f0 = long(cx) * s + ex.first, f1 = long(cy) * s + ey.first;
              const long   f2 = dim == 3 ? long(cz) * s + ez.first : 0;
              const double w  = ex.second * ey.second * (dim == 3 ? ez.second : 1.0);
              const double kv = kappa[f0 + (long)fine_nc[0] * (f1 + (long)fine_nc[1] * f2)];
              acc += w * (mode == 1 ? std::log(kv) : mode == 2 ? 1.0 / kv : kv);
              wsum += w;
            }
      acc /= wsum;
      return mode == 1 ? std::exp(acc) : mode == 2 ? 1.0 / acc : acc;
    }
  };

  // Element mass and stiffness matrices of cell c (lexicographic local dofs).
  static void element_matrices(const Space &sp, const Shape1D &sh, const Coefficient &coef,
                               int lev_shift, long c, std::vector<double> &Me,
                               std::vector<double> &Ke)
  {
    const int dim = sp.dim, n = sp.n, nq = (int)sh.gq.size();
    const int nl = sp.nc_local;
    Me.assign(size_t(nl) * nl, 0.0);
    Ke.assign(size_t(nl) * nl, 0.0);
    double V[8][3];
    sp.cell_vertices(c, V);
    int cx, cy, cz;
    sp.cell_coords(c, cx, cy, cz);
    std::vector<double> phi(nl), gr(size_t(nl) * 3);
    for (int qz = 0; qz < (dim == 3 ? nq : 1); ++qz)
      for (int qy = 0; qy < nq; ++qy)
        for (int qx = 0; qx < nq; ++qx)
          {
            double xi[3] = {sh.gq[qx], sh.gq[qy], dim == 3 ? sh.gq[qz] : 0.0};
            double w     = sh.gw[qx] * sh.gw[qy] * (dim == 3 ? sh.gw[qz] : 1.0);
            double x[3], J[3][3], Ji[3][3];
            q1_map(dim, V, xi, x, J);
            double det = det_inv(dim, J, Ji);
            double JxW = std::abs(det) * w;
            double kap = coef.at(lev_shift, cx, cy, cz, xi);
            int    l   = 0;
            for (int c3 = 0; c3 < (dim == 3 ? n : 1); ++c3)
              for (int b = 0; b < n; ++b)
                for (int a = 0; a < n; ++a, ++l)
                  {
                    double sx = sh.S[qx * n + a], sy = sh.S[qy * n + b];
                    double sz = (dim == 3) ? sh.S[qz * n + c3] : 1.0;
                    double dx = sh.D[qx * n + a], dy = sh.D[qy * n + b];
                    double dz = (dim == 3) ? sh.D[qz * n + c3] : 0.0;
                    phi[l]    = sx * sy * sz;
                    double g[3] = {dx * sy * sz, sx * dy * sz, sx * sy * dz};
                    for (int d = 0; d < dim; ++d)
                      {
                        double s = 0;
                        for (int e = 0; e < dim; ++e)
                          s += Ji[e][d] * g[e];
                        gr[l * 3 + d] = s;
                      }
                  }
            for (int i = 0; i < nl; ++i)
              for (int j = 0; j < nl; ++j)
                {
                  Me[size_t(i) * nl + j] += phi[i] * phi[j] * JxW;
                  double s = 0;
                  for (int d = 0; d < dim; ++d)
                    s += gr[i * 3 + d] * gr[j * 3 + d];
                  Ke[size_t(i) * nl + j] += kap * s * JxW;
                }
          }
  }

  // Assembled K and M with the constrained-dof convention of
  // AffineConstraints::distribute_local_to_global.
  static void assemble(const Space &sp, const Shape1D &sh, const Coefficient &coef,
                       int lev_shift, CSR &K, CSR &M)
  {
    const long nd = sp.ndofs;
    const int  nl = sp.nc_local;
    std::vector<std::vector<long>> cols(nd);
    std::vector<long>              idx;
    for (long c = 0; c < sp.g.n_cells(); ++c)
      {
        sp.cell_dofs(c, idx);
        for (int i = 0; i < nl; ++i)
          for (int j = 0; j < nl; ++j)
            cols[idx[i]].push_back(idx[j]);
      }
    K.n = K.m = M.n = M.m = nd;
    K.rp.assign(nd + 1, 0);
    for (long i = 0; i < nd; ++i)
      {
        auto &c = cols[i];
        std::sort(c.begin(), c.end());
        c.erase(std::unique(c.begin(), c.end()), c.end());
        K.rp[i + 1] = K.rp[i] + (long)c.size();
      }
    K.ci.resize(K.rp[nd]);
    for (long i = 0; i < nd; ++i)
      std::copy(cols[i].begin(), cols[i].end(), K.ci.begin() + K.rp[i]);
    K.v.assign(K.ci.size(), 0.0);
    M.rp = K.rp;
    M.ci = K.ci;
    M.v.assign(K.ci.size(), 0.0);
    auto find = [&](long i, long j) {
      auto b = K.ci.begin() + K.rp[i], e = K.ci.begin() + K.rp[i + 1];
      return long(std::lower_bound(b, e, j) - K.ci.begin());
    };
    std::vector<double> Me, Ke;
    for (long c = 0; c < sp.g.n_cells(); ++c)
      {
        sp.cell_dofs(c, idx);
        element_matrices(sp, sh, coef, lev_shift, c, Me, Ke);
        for (int i = 0; i < nl; ++i)
          {
            const long gi = idx[i];
            if (sp.constrained[gi])
              {
                long pos = find(gi, gi);
                K.v[pos] += std::abs(Ke[size_t(i) * nl + i]);
                M.v[pos] += std::abs(Me[size_t(i) * nl + i]);
                continue;
              }
            for (int j = 0; j < nl; ++j)
              {
                const long gj = idx[j];
                if (sp.constrained[gj])
                  continue;
                long pos = find(gi, gj);
                K.v[pos] += Ke[size_t(i) * nl + j];
                M.v[pos] += Me[size_t(i) * nl + j];
              }
          }
      }
  }

  // y = A_ff x on free dofs (constrained rows zero, constrained columns
  // ignored): the matrix-free operator's semantics.
  template <typename T>
  static void spmv_free(const CSR &A, const std::vector<T> &vals, const std::vector<char> &cons,
                        const T *x, T *y)
  {
    // rows are independent: threads change nothing in the arithmetic
#pragma omp parallel for schedule(static)
    for (long i = 0; i < A.n; ++i)
      {
        if (cons[i])
          {
            y[i] = T(0);
            continue;
          }
        T s = T(0);
        for (long k = A.rp[i]; k < A.rp[i + 1]; ++k)
          {
            long j = A.ci[k];
            if (!cons[j])
              s += vals[k] * x[j];
          }
        y[i] = s;
      }
  }


  // deal.II DoF enumeration of FE_Q(p) on a globally refined single 2D
  // coarse cell: active cells in hierarchical (Morton) order; per cell the
  // unnumbered vertices (lexicographic), then lines 0..3 (x=0, x=1, y=0,
  // y=1; dofs in increasing coordinate), then the interior (lexicographic).
  // Returns dealii_index[lexicographic index].  Used only for the power
  // iteration start vector, the one place where the numbering is visible.
  static std::vector<long> dealii_numbering_2d(const Space &sp)
  {
    const int  p = sp.p, ncx = sp.g.nc[0];
    int        levels = 0;
    while ((1 << levels) < ncx)
      ++levels;
    std::vector<long> num(sp.ndofs, -1);
    long              next = 0;
    auto gid = [&](long ix, long iy) { return ix + sp.N[0] * iy; };
    for (long m = 0; m < (long)ncx * ncx; ++m)
      {
        long cx = 0, cy = 0;
        for (int b = 0; b < levels; ++b)
          {
            cx |= ((m >> (2 * b)) & 1) << b;
            cy |= ((m >> (2 * b + 1)) & 1) << b;
          }
        const long x0 = cx * p, y0 = cy * p;
        const long vx[4] = {x0, x0 + p, x0, x0 + p}, vy[4] = {y0, y0, y0 + p, y0 + p};
        for (int v = 0; v < 4; ++v)
          if (num[gid(vx[v], vy[v])] < 0)
            num[gid(vx[v], vy[v])] = next++;
        // lines: 0: x=x0 along y, 1: x=x0+p along y, 2: y=y0 along x, 3: y=y0+p along x
        for (int l = 0; l < 4; ++l)
          {
            if (p < 2)
              break;
            long first = (l < 2) ? gid((l == 0) ? x0 : x0 + p, y0 + 1) :
                                   gid(x0 + 1, (l == 2) ? y0 : y0 + p);
            if (num[first] >= 0)
              continue;
            for (int k = 1; k < p; ++k)
              {
                long g = (l < 2) ? gid((l == 0) ? x0 : x0 + p, y0 + k) :
                                   gid(x0 + k, (l == 2) ? y0 : y0 + p);
                num[g] = next++;
              }
          }
        for (int j = 1; j < p; ++j)
          for (int i = 1; i < p; ++i)
            num[gid(x0 + i, y0 + j)] = next++;
      }
    return num;
  }

  // ------------------------------------------------------------- levels
  template <typename T>
  struct Level
  {
    Space               sp;
    CSR                 K, M;         // assembled, FP64 values (Dirichlet diag)
    std::vector<T>      Kv, Mv;       // operator values in T
    int                 nt = 1;
    Mat                 Alpha, Beta;  // FP64
    std::vector<T>      At, Bt;       // in T
    std::vector<std::vector<T>> Binv; // Vanka inverses per cell
    std::vector<T>      inv_valence;
    double              omega = 1.0;
    int                 r = 0, nts = 1;
    long                N() const { return sp.ndofs; }
    long                size() const { return sp.ndofs * nt; }

    // SystemMatrix::vmult (operators.h:94-122), block-major vectors.
    void vmult(T *dst, const T *src, bool transpose = false) const
    {
      const long     n = N();
      std::vector<T> tmp(n);
      std::fill(dst, dst + n * nt, T(0));
      for (int i = 0; i < nt; ++i)
        {
          spmv_free(K, Kv, sp.constrained, src + i * n, tmp.data());
          for (int j = 0; j < nt; ++j)
            {
              T a = transpose ? At[i * nt + j] : At[j * nt + i];
              if (a != T(0))
                for (long k = 0; k < n; ++k)
                  dst[j * n + k] += a * tmp[k];
            }
        }
      for (int i = 0; i < nt; ++i)
        {
          spmv_free(M, Mv, sp.constrained, src + i * n, tmp.data());
          for (int j = 0; j < nt; ++j)
            {
              T b = transpose ? Bt[i * nt + j] : Bt[j * nt + i];
              if (b != T(0))
                for (long k = 0; k < n; ++k)
                  dst[j * n + k] += b * tmp[k];
            }
        }
    }

    // PreconditionVanka ctor (gmg.h:409-474).  Inverses formed in precision
    // TI: the reference inverts in the multigrid number type.
    void setup_vanka(bool invert_in_T)
    {
      const int  nl = sp.nc_local;
      const long nc = sp.g.n_cells();
      std::vector<double> val(N(), 0.0);
      Binv.assign(nc, {});
      const int nb = nt * nl;
      {
        std::vector<long> idx;
        for (long c = 0; c < nc; ++c)
          {
            sp.cell_dofs(c, idx);
            for (long g : idx)
              val[g] += 1.0;
          }
      }
      // the cells' blocks are independent (one thread per cell)
      bool singular = false;
#pragma omp parallel for schedule(dynamic)
      for (long c = 0; c < nc; ++c)
        {
          std::vector<long> idx;
          sp.cell_dofs(c, idx);
          std::vector<double> Kc(size_t(nl) * nl, 0.0), Mc(size_t(nl) * nl, 0.0);
          for (int i = 0; i < nl; ++i)
            {
              long gi = idx[i];
              for (int j = 0; j < nl; ++j)
                {
                  long gj = idx[j];
                  auto b = K.ci.begin() + K.rp[gi], e = K.ci.begin() + K.rp[gi + 1];
                  auto it = std::lower_bound(b, e, gj);
                  if (it != e && *it == gj)
                    {
                      long pos = it - K.ci.begin();
                      // the reference restricts its MG-precision matrices
                      Kc[size_t(i) * nl + j] = invert_in_T ? double(T(K.v[pos])) : K.v[pos];
                      Mc[size_t(i) * nl + j] = invert_in_T ? double(T(M.v[pos])) : M.v[pos];
                    }
                }
            }
          if (invert_in_T)
            {
              std::vector<T> B(size_t(nb) * nb, T(0));
              for (int i = 0; i < nt; ++i)
                for (int j = 0; j < nt; ++j)
                  if (Beta(i, j) != 0.0 || Alpha(i, j) != 0.0)
                    for (int k = 0; k < nl; ++k)
                      for (int l = 0; l < nl; ++l)
                        B[size_t(k + i * nl) * nb + l + j * nl] =
                          Bt[i * nt + j] * T(Mc[size_t(k) * nl + l]) +
                          At[i * nt + j] * T(Kc[size_t(k) * nl + l]);
              if (!invert_inplace(B, nb))
                {
#pragma omp atomic write
                  singular = true;
                }
              Binv[c] = B;
            }
          else
            {
              std::vector<double> B(size_t(nb) * nb, 0.0);
              for (int i = 0; i < nt; ++i)
                for (int j = 0; j < nt; ++j)
                  if (Beta(i, j) != 0.0 || Alpha(i, j) != 0.0)
                    for (int k = 0; k < nl; ++k)
                      for (int l = 0; l < nl; ++l)
                        B[size_t(k + i * nl) * nb + l + j * nl] =
                          Beta(i, j) * Mc[size_t(k) * nl + l] + Alpha(i, j) * Kc[size_t(k) * nl + l];
              if (!invert_inplace(B, nb))
                {
#pragma omp atomic write
                  singular = true;
                }
              Binv[c].assign(B.begin(), B.end());
            }
        }
      if (singular)
        {
          fprintf(stderr, "oracle: singular Vanka block\n");
          abort();
        }
      inv_valence.resize(N());
      for (long i = 0; i < N(); ++i)
        inv_valence[i] = T(1.0 / val[i]);
    }

    // PreconditionVanka::vmult (gmg.h:476-514)
    void vanka(T *dst, const T *src) const
    {
      const int  nl = sp.nc_local, nb = nt * nl;
      const long n  = N();
      const long nc = sp.g.n_cells();
      std::fill(dst, dst + n * nt, T(0));
      // local solves in parallel, scatter in cell order (the same sums as a
      // serial loop whatever the thread count)
      std::vector<T> Y(size_t(nc) * nb);
#pragma omp parallel
      {
        std::vector<long> idx;
        std::vector<T>    xl(nb);
#pragma omp for schedule(static)
        for (long c = 0; c < nc; ++c)
          {
            sp.cell_dofs(c, idx);
            for (int b = 0, k = 0; b < nt; ++b)
              for (int j = 0; j < nl; ++j, ++k)
                xl[k] = src[b * n + idx[j]];
            const auto &B  = Binv[c];
            T          *yl = Y.data() + size_t(c) * nb;
            for (int i = 0; i < nb; ++i)
              {
                T s = T(0);
                for (int j = 0; j < nb; ++j)
                  s += B[size_t(i) * nb + j] * xl[j];
                yl[i] = s;
              }
          }
      }
      std::vector<long> idx;
      for (long c = 0; c < nc; ++c)
        {
          sp.cell_dofs(c, idx);
          const T *yl = Y.data() + size_t(c) * nb;
          for (int b = 0, k = 0; b < nt; ++b)
            for (int j = 0; j < nl; ++j, ++k)
              dst[b * n + idx[j]] += inv_valence[idx[j]] * yl[k];
        }
    }
  };

  // Space (h) transfer between nested structured levels: P(fine, coarse)
  // = coarse basis evaluated at the fine node, zero Dirichlet both sides.
  static CSR space_prolongation(const Space &f, const Space &c, const Shape1D &sh)
  {
    CSR P;
    P.n = f.ndofs;
    P.m = c.ndofs;
    P.rp.assign(P.n + 1, 0);
    const int dim = f.dim, p = f.p, n = f.n;
    // per direction: for fine node index i, coarse cell cc and 1D weights
    for (long k = 0; k < f.N[2]; ++k)
      for (long j = 0; j < f.N[1]; ++j)
        for (long i = 0; i < f.N[0]; ++i)
          {
            long fi = i + f.N[0] * (j + f.N[1] * k);
            if (!f.constrained[fi])
              {
                long   fidx[3] = {i, j, k};
                int    cc[3];
                double w1[3][8];
                for (int d = 0; d < 3; ++d)
                  {
                    if (d >= dim)
                      {
                        cc[d] = 0;
                        continue;
                      }
                    long fc  = fidx[d] / p; // fine cell
                    int  a   = int(fidx[d] % p);
                    if (fc >= f.g.nc[d])
                      fc = f.g.nc[d] - 1, a = p;
                    cc[d]        = int(fc / 2);
                    int    child = int(fc % 2);
                    double xi    = 0.5 * (sh.gll[a] + child);
                    for (int b = 0; b < n; ++b)
                      w1[d][b] = sh.L.value(b, xi);
                  }
                for (int c3 = 0; c3 < (dim == 3 ? n : 1); ++c3)
                  for (int b = 0; b < n; ++b)
                    for (int a = 0; a < n; ++a)
                      {
                        double w = w1[0][a] * w1[1][b] * (dim == 3 ? w1[2][c3] : 1.0);
                        if (std::abs(w) < 1e-15)
                          continue;
                        long ci = (cc[0] * p + a) +
                                  c.N[0] * ((cc[1] * p + b) + c.N[1] * (long)(cc[2] * p + c3));
                        if (c.constrained[ci])
                          continue;
                        P.ci.push_back(ci);
                        P.v.push_back(w);
                      }
              }
            P.rp[fi + 1] = (long)P.ci.size();
          }
    return P;
  }

  template <typename T>
  struct Transfer
  {
    int            kind = MG_NONE; // MG_NONE: space, else time
    CSR            P;              // space prolongation
    std::vector<T> Pv;
    Mat            Pt, Rt;         // time prolongation / restriction
    std::vector<T> Ptt, Rtt;
    int            nt_f = 1, nt_c = 1;
    long           nf = 0, nc = 0;
    bool           overwrite = false; // transfer_01.cc:42-83 assigns instead of adding

    void prolongate_and_add(T *dst, const T *src) const
    {
      if (kind == MG_NONE)
        {
          for (int b = 0; b < nt_f; ++b)
            for (long i = 0; i < nf; ++i)
              {
                T s = T(0);
                for (long k = P.rp[i]; k < P.rp[i + 1]; ++k)
                  s += Pv[k] * src[b * nc + P.ci[k]];
                dst[b * nf + i] += s;
              }
        }
      else
        for (long i = 0; i < nf; ++i)
          for (int a = 0; a < nt_f; ++a)
            {
              T s = T(0);
              for (int b = 0; b < nt_c; ++b)
                if (Pt(a, b) != 0.0)
                  s += Ptt[a * nt_c + b] * src[b * nc + i];
              if (overwrite)
                dst[a * nf + i] = s;
              else
                dst[a * nf + i] += s;
            }
    }
    void restrict_and_add(T *dst, const T *src) const
    {
      if (kind == MG_NONE)
        {
          for (int b = 0; b < nt_f; ++b)
            for (long i = 0; i < nf; ++i)
              for (long k = P.rp[i]; k < P.rp[i + 1]; ++k)
                dst[b * nc + P.ci[k]] += Pv[k] * src[b * nf + i];
        }
      else
        for (long i = 0; i < nf; ++i)
          for (int a = 0; a < nt_c; ++a)
            {
              T s = T(0);
              for (int b = 0; b < nt_f; ++b)
                if (Rt(a, b) != 0.0)
                  s += Rtt[a * nt_f + b] * src[b * nf + i];
              if (overwrite)
                dst[a * nc + i] = s;
              else
                dst[a * nc + i] += s;
            }
    }
  };

  // ---------------------------------------------------------------- spec
  struct Spec
  {
    int                 dim = 2;
    int                 subdiv[3] = {1, 1, 1};
    double              lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    int                 refine = 2;
    int                 p = -1; // space degree (default r+1 as tp_01.cc:77)
    int                 type = DG;
    int                 r = 1;
    int                 nts = 1;
    double              tau = -1; // default tp_01.cc:124-127 rule
    int                 space_time_mg = 1;
    int                 nts_min = -1, r_min = -1;
    int                 time_before_space = 0;
    int                 mg_space = 1;     // 0: all levels on the fine mesh (transfer_01)
    int                 mg_float = 0;
    int                 vanka_invert_T = 1;
    int                 smoothing_steps = 1;
    int                 variable = 1;
    int                 estimate_relaxation = 1;
    int                 overwrite_transfer = 0;
    double              eig_safety = 1.15;
    int                 omega_rule = 0; // 0: alpha=min(0.9 lmax, lmin=+inf) ; 1: alpha = 0
    int                 eig_iters = 20;
    double              reltol = 1e-12, abstol = 1e-12;
    int                 maxiter = 200, basis = 30;
    int                 extrapolate = 1;
    double              end_time = 1.0;
    double              frequency = 1.0;
    int                 rhs = 0; // 0: manufactured solution's source, 1: zero, 3: Gaussian bumps
    int                 n_bumps = 4;  // rhs=3: number of bumps
    long                seed = 1;     // rhs=3: generator seed
    int                 r_is_pT = 1;  // time restriction = P^T (gmg.h:557)
    int                 problem = 0; // 0: heat, 1: wave (tp_01.cc:167-197, time_integrators.h:183-280)
    int                 coarse_coef = 0; // coarse-level coefficient: 0 sampled, 1/2/3 geometric/harmonic/arithmetic mean
    int                 dirichlet = 63;  // zero-Dirichlet faces (bit 2d: low side, 2d+1: high side)
    int                 transpose_time = 0; // every level's time matrices transposed (adjoint system)
    int                 fgmres_variant = 0; // 0: deal.II 9.5 projected solve at j>0
    int                 dealii_order = 0;   // power-iteration start vector in deal.II dof order
    double              omega_scale = 1.0;  // diagnostic multiplier on estimated omegas
    int                 start_ones = 1;     // start vector: ones with entry 0 zeroed
    int                 est_mode = 0;       // 0: Rayleigh quotient, 1: norm of P^-1 A v
  };

  static Spec parse_spec(const char *s)
  {
    Spec               sp;
    std::string        str(s ? s : "");
    std::stringstream  ss(str);
    std::string        tok;
    std::map<std::string, std::string> kv;
    while (std::getline(ss, tok, ';'))
      {
        auto e = tok.find('=');
        if (e == std::string::npos)
          continue;
        kv[tok.substr(0, e)] = tok.substr(e + 1);
      }
    auto geti = [&](const char *k, int &v) {
      if (kv.count(k))
        v = std::atoi(kv[k].c_str());
    };
    auto getd = [&](const char *k, double &v) {
      if (kv.count(k))
        v = std::atof(kv[k].c_str());
    };
    geti("dim", sp.dim);
    if (kv.count("subdiv"))
      {
        std::stringstream s2(kv["subdiv"]);
        for (int d = 0; d < 3 && std::getline(s2, tok, ','); ++d)
          sp.subdiv[d] = std::atoi(tok.c_str());
      }
    geti("refine", sp.refine);
    geti("p", sp.p);
    if (kv.count("type"))
      sp.type = (kv["type"] == "CGP" || kv["type"] == "1") ? CGP : DG;
    geti("r", sp.r);
    geti("nts", sp.nts);
    getd("tau", sp.tau);
    geti("space_time_mg", sp.space_time_mg);
    geti("nts_min", sp.nts_min);
    geti("r_min", sp.r_min);
    geti("time_before_space", sp.time_before_space);
    geti("mg_space", sp.mg_space);
    geti("mg_float", sp.mg_float);
    geti("vanka_invert_T", sp.vanka_invert_T);
    geti("smoothing_steps", sp.smoothing_steps);
    geti("variable", sp.variable);
    geti("estimate_relaxation", sp.estimate_relaxation);
    geti("overwrite_transfer", sp.overwrite_transfer);
    getd("eig_safety", sp.eig_safety);
    geti("omega_rule", sp.omega_rule);
    geti("eig_iters", sp.eig_iters);
    getd("reltol", sp.reltol);
    getd("abstol", sp.abstol);
    geti("maxiter", sp.maxiter);
    geti("basis", sp.basis);
    geti("extrapolate", sp.extrapolate);
    getd("end_time", sp.end_time);
    getd("frequency", sp.frequency);
    geti("rhs", sp.rhs);
    geti("n_bumps", sp.n_bumps);
    if (kv.count("seed"))
      sp.seed = std::atol(kv["seed"].c_str());
    geti("r_is_pT", sp.r_is_pT);
    if (kv.count("problem"))
      sp.problem = (kv["problem"] == "wave" || kv["problem"] == "1") ? 1 : 0;
    geti("coarse_coef", sp.coarse_coef);
    geti("dirichlet", sp.dirichlet);
    geti("transpose_time", sp.transpose_time);
    geti("fgmres_variant", sp.fgmres_variant);
    geti("dealii_order", sp.dealii_order);
    getd("omega_scale", sp.omega_scale);
    geti("start_ones", sp.start_ones);
    geti("est_mode", sp.est_mode);
    // Parameters::parse derivations (gmg.h:637-648)
    if (sp.nts_min == -1)
      sp.nts_min = sp.nts / 2;
    sp.nts_min = std::clamp(sp.nts_min, 1, sp.nts);
    const int lowest = sp.type == DG ? 0 : 1;
    if (sp.r_min == -1)
      sp.r_min = sp.r - 1;
    sp.r_min = std::clamp(sp.r_min, lowest, sp.r);
    if (sp.p < 0)
      sp.p = sp.r + 1;
    return sp;
  }

  // ---------------------------------------------------------- the problem
  template <typename T>
  struct Problem
  {
    Spec                       spec;
    Shape1D                    sh;
    Coefficient                coef;
    Grid                       fine_grid;
    std::vector<int>           mg_type;
    std::vector<Level<T>>      lv;       // MG levels in precision T
    std::vector<Transfer<T>>   tr;       // tr[l]: between l-1 and l
    Level<double>              fine;     // FP64 system matrix (fine level)
    Mat                        rhsK, rhsM; // rhs_matrix weights
    Mat                        Alpha1, Gamma1;
    // wave: rhs_matrix_v weights (mass only) and the velocity recovery
    // matrices of TimeIntegratorWave (time_integrators.h:219-233)
    Mat                        rhsV, AixB, AixG, AixZ;
    int                        nt = 1;
    double                     tau = 0.0;
    std::vector<double>        omega_log;
    std::vector<double>        lambda_log; // raw power-iteration estimates per level
    // FGMRES statistics of the last solve
    int                        last_its = 0;
    std::vector<double>        res_hist;

    Grid grid_at(int s) const
    {
      // refinement level s in [0, refine]; coarse vertices are a subset of
      // the finest vertices.
      Grid g;
      g.dim         = spec.dim;
      g.dirichlet   = spec.dirichlet;
      const int fac = 1 << (spec.refine - s);
      for (int d = 0; d < 3; ++d)
        g.nc[d] = (d < spec.dim) ? fine_grid.nc[d] / fac : 1;
      const long nvx = g.nc[0] + 1, nvy = g.nc[1] + 1, nvz = (spec.dim == 3) ? g.nc[2] + 1 : 1;
      g.X.resize(size_t(nvx * nvy * nvz) * spec.dim);
      for (long k = 0; k < nvz; ++k)
        for (long j = 0; j < nvy; ++j)
          for (long i = 0; i < nvx; ++i)
            {
              long fid = fine_grid.vid(int(i * fac), int(j * fac), int(k * fac));
              long id  = i + nvx * (j + nvy * k);
              for (int d = 0; d < spec.dim; ++d)
                g.X[id * spec.dim + d] = fine_grid.X[fid * spec.dim + d];
            }
      return g;
    }

    void build_level(Level<T> &L, const Grid &g, int lev_shift, int type, int r, int nts,
                     double tau_l)
    {
      L.sp.init(g, spec.p);
      assemble(L.sp, sh, coef, lev_shift, L.K, L.M);
      L.Kv.assign(L.K.v.begin(), L.K.v.end());
      L.Mv.assign(L.M.v.begin(), L.M.v.end());
      auto w    = fe_time_weights(type, r, tau_l, nts);
      if (spec.problem == 1)
        {
          // level weights of the wave problem: the Schur-in-time transform
          // of the level's (multi-step) weights taken as one interval
          // (fe_time.h:383-412, which calls the transform with its default
          // n_timesteps_at_once = 1)
          auto ww = fe_time_weights_wave(type, w[0], w[1], w[2], w[3], 1);
          w[0]    = ww[0];
          w[1]    = ww[1];
        }
      L.Alpha   = spec.transpose_time ? transpose(w[0]) : w[0];
      L.Beta    = spec.transpose_time ? transpose(w[1]) : w[1];
      L.nt      = w[0].m;
      L.r       = r;
      L.nts     = nts;
      L.At.assign(L.Alpha.a.begin(), L.Alpha.a.end());
      L.Bt.assign(L.Beta.a.begin(), L.Beta.a.end());
    }

    void init(const Spec &s, const double *coords, const double *kappa)
    {
      spec = s;
      if (spec.rhs == 3)
        make_bumps();
      sh.init(spec.p, spec.p + 1);
      // fine grid
      fine_grid.dim       = spec.dim;
      fine_grid.dirichlet = spec.dirichlet;
      for (int d = 0; d < 3; ++d)
        fine_grid.nc[d] = (d < spec.dim) ? spec.subdiv[d] << spec.refine : 1;
      const long nvx = fine_grid.nc[0] + 1, nvy = fine_grid.nc[1] + 1,
                 nvz = (spec.dim == 3) ? fine_grid.nc[2] + 1 : 1;
      fine_grid.X.resize(size_t(nvx * nvy * nvz) * spec.dim);
      for (long k = 0; k < nvz; ++k)
        for (long j = 0; j < nvy; ++j)
          for (long i = 0; i < nvx; ++i)
            {
              long   id = i + nvx * (j + nvy * k);
              long   ii[3] = {i, j, k};
              for (int d = 0; d < spec.dim; ++d)
                fine_grid.X[id * spec.dim + d] =
                  coords ? coords[id * spec.dim + d] :
                           spec.lo[d] + (spec.hi[d] - spec.lo[d]) * double(ii[d]) / fine_grid.nc[d];
            }
      coef.dim = spec.dim;
      for (int d = 0; d < 3; ++d)
        coef.fine_nc[d] = fine_grid.nc[d];
      if (kappa)
        coef.kappa.assign(kappa, kappa + fine_grid.n_cells());
      coef.mode = spec.coarse_coef;
      coef.gq   = sh.gq;
      // tau rule (tp_01.cc:87, 124-127)
      if (spec.tau <= 0)
        {
          double dmin = 1e300;
          for (int d = 0; d < spec.dim; ++d)
            (void)d;
          // coarse cells of a subdivided box: diameter of one coarse cell
          double diag = 0;
          for (int d = 0; d < spec.dim; ++d)
            {
              double h = (spec.hi[d] - spec.lo[d]) / spec.subdiv[d];
              diag += h * h;
            }
          dmin            = std::sqrt(diag);
          double spc_step = dmin / std::sqrt((double)spec.dim);
          unsigned nsteps = (unsigned)(spec.end_time / spc_step);
          tau             = spec.end_time * std::pow(2.0, -(spec.refine + 1)) / nsteps;
        }
      else
        tau = spec.tau;

      // MG level types (tp_01.cc:205-219)
      const int n_sp = spec.mg_space ? spec.refine + 1 : 1;
      const int rmin = spec.space_time_mg ? spec.r_min : spec.r;
      const int tmin = spec.space_time_mg ? std::max(spec.nts_min, 1) : spec.nts;
      mg_type = time_mg_sequence(n_sp, spec.r, rmin, spec.nts, tmin, spec.time_before_space);
      const int nlev = (int)mg_type.size() + 1;
      // space refinement index, r, nts, tau per level
      std::vector<int>    sref(nlev), rl(nlev), nl(nlev);
      std::vector<double> tl(nlev);
      sref[nlev - 1] = spec.mg_space ? spec.refine : spec.refine;
      rl[nlev - 1]   = spec.r;
      nl[nlev - 1]   = spec.nts;
      tl[nlev - 1]   = tau;
      for (int l = nlev - 2; l >= 0; --l)
        {
          int t   = mg_type[l];
          sref[l] = sref[l + 1] - ((t == MG_NONE && spec.mg_space) ? 1 : 0);
          rl[l]   = rl[l + 1] - (t == MG_K ? 1 : 0);
          nl[l]   = nl[l + 1] / (t == MG_TAU ? 2 : 1);
          tl[l]   = tl[l + 1] * (t == MG_TAU ? 2.0 : 1.0);
        }
      lv.resize(nlev);
      for (int l = 0; l < nlev; ++l)
        {
          build_level(lv[l], grid_at(sref[l]), spec.refine - sref[l], spec.type, rl[l], nl[l],
                      tl[l]);
          lv[l].setup_vanka(spec.vanka_invert_T != 0);
        }
      tr.resize(nlev);
      for (int l = 1; l < nlev; ++l)
        {
          auto &t  = tr[l];
          t.kind   = mg_type[l - 1];
          t.nt_f   = lv[l].nt;
          t.nt_c   = lv[l - 1].nt;
          t.nf     = lv[l].N();
          t.nc     = lv[l - 1].N();
          t.overwrite = spec.overwrite_transfer != 0;
          if (t.kind == MG_NONE)
            {
              t.P = space_prolongation(lv[l].sp, lv[l - 1].sp, sh);
              t.Pv.assign(t.P.v.begin(), t.P.v.end());
            }
          else
            {
              // gmg.h:86-115: R = P^T (restrict_is_transpose_prolongate, the
              // default of gmg.h:557) or the restriction / projection matrix
              t.Pt = (t.kind == MG_K) ? time_projection(spec.type, rl[l] - 1, rl[l], nl[l]) :
                                        time_prolongation(spec.type, rl[l], nl[l]);
              if (spec.r_is_pT)
                t.Rt = transpose(t.Pt);
              else
                t.Rt = (t.kind == MG_K) ? time_projection(spec.type, rl[l], rl[l] - 1, nl[l]) :
                                          time_restriction(spec.type, rl[l], nl[l]);
              t.Ptt.assign(t.Pt.a.begin(), t.Pt.a.end());
              t.Rtt.assign(t.Rt.a.begin(), t.Rt.a.end());
            }
        }
      // fine FP64 system
      {
        Level<double> &F = fine;
        F.sp.init(fine_grid, spec.p);
        assemble(F.sp, sh, coef, 0, F.K, F.M);
        F.Kv = F.K.v;
        F.Mv = F.M.v;
        auto w   = fe_time_weights(spec.type, spec.r, tau, spec.nts);
        auto w1  = fe_time_weights(spec.type, spec.r, tau, 1);
        Mat  zero(w[2].m, 1);
        if (spec.problem == 1)
          {
            // tp_01.cc:167-190: lhs = (Alpha_lhs, Beta_lhs), rhs_matrix =
            // (rhs_uK, rhs_uM) on prev_u, rhs_matrix_v = (0, rhs_vM) on prev_v
            auto ww = fe_time_weights_wave(spec.type, w1[0], w1[1], w1[2], w1[3], spec.nts);
            w[0]    = ww[0];
            w[1]    = ww[1];
            rhsK    = ww[2];
            rhsM    = ww[3];
            rhsV    = ww[4];
            // TimeIntegratorWave: AixB = Alpha^-1 Beta, AixG = Alpha^-1 Gamma
            // (negated for dG), AixZ = Alpha^-1 Zeta (negated for cG)
            const Mat Ai = inverse(w1[0]);
            AixB         = matmul(Ai, w1[1]);
            AixG         = matmul(Ai, w1[2]);
            AixZ         = matmul(Ai, w1[3]);
            if (spec.type == DG)
              for (auto &e : AixG.a)
                e = -e;
            else
              for (auto &e : AixZ.a)
                e = -e;
          }
        else if (spec.type == CGP)
          rhsK = w[2], rhsM = w[3];
        else
          rhsK = zero, rhsM = w[2];
        F.Alpha  = spec.transpose_time ? transpose(w[0]) : w[0];
        F.Beta   = spec.transpose_time ? transpose(w[1]) : w[1];
        F.nt     = w[0].m;
        F.At     = F.Alpha.a;
        F.Bt     = F.Beta.a;
        nt       = F.nt;
        Alpha1  = w1[0];
        Gamma1  = w1[2];
      }
      // relaxation parameters (gmg.h:708-844)
      omega_log.assign(nlev, 1.0);
      lambda_log.assign(nlev, 0.0);
      for (int l = 0; l < nlev; ++l)
        {
          lv[l].omega  = spec.estimate_relaxation ? estimate_omega(l) : 1.0;
          omega_log[l] = lv[l].omega;
        }
    }

    // PreconditionChebyshev::estimate_eigenvalues, power iteration branch.
    double estimate_omega(int l)
    {
      if (l == 0)
        return 1.0;
      auto          &L = lv[l];
      const long     n = L.N(), sz = L.size();
      std::vector<T> ev(sz), v1(sz), v2(sz);
      std::vector<long> order(n);
      std::iota(order.begin(), order.end(), 0L);
      if (spec.dealii_order && L.sp.dim == 2 && L.sp.g.nc[0] == L.sp.g.nc[1] &&
          (L.sp.g.nc[0] & (L.sp.g.nc[0] - 1)) == 0)
        order = dealii_numbering_2d(L.sp);
      for (int b = 0; b < L.nt; ++b)
        {
          double mean = 0;
          for (long i = 0; i < n; ++i)
            mean += double(order[i] % 11);
          mean /= n;
          for (long i = 0; i < n; ++i)
            ev[b * n + i] = T(double(order[i] % 11) - mean);
          if (spec.start_ones)
            for (long i = 0; i < n; ++i)
              ev[b * n + i] = T((b == 0 && order[i] == 0) ? 0.0 : 1.0);
        }
      auto nrm = [&](const std::vector<T> &v) {
        double s = 0;
        for (T x : v)
          s += double(x) * double(x);
        return std::sqrt(s);
      };
      double en = nrm(ev);
      for (auto &x : ev)
        x = T(x / en);
      double est = 0;
      for (int it = 0; it < spec.eig_iters; ++it)
        {
          if (spec.est_mode == 2)
            {
              L.vanka(v2.data(), ev.data());
              L.vmult(v1.data(), v2.data());
            }
          else
            {
              L.vmult(v2.data(), ev.data());
              L.vanka(v1.data(), v2.data());
            }
          double d = 0;
          for (long i = 0; i < sz; ++i)
            d += double(ev[i]) * double(v1[i]);
          double n1 = nrm(v1);
          est       = spec.est_mode != 1 ? d : n1;
          for (long i = 0; i < sz; ++i)
            v1[i] = T(v1[i] / n1);
          ev.swap(v1);
        }
      // a level without free unknowns (every node on a Dirichlet face) has
      // nothing to estimate
      if (!std::isfinite(est) || est == 0.0)
        return 1.0;
      lambda_log[l] = est;
      double lmax  = std::abs(est) * spec.eig_safety;
      double alpha = spec.omega_rule == 0 ? 0.9 * lmax : 0.0;
      return spec.omega_scale * 2.0 / (alpha + lmax);
    }

    // PreconditionRelaxation step_operations with PreconditionVanka
    void relax(int l, T *u, const T *rhs, bool first) const
    {
      const auto &L  = lv[l];
      const long  sz = L.size();
      const T     om = T(L.omega);
      if (first)
        {
          L.vanka(u, rhs);
          if (L.omega != 1.0)
            for (long i = 0; i < sz; ++i)
              u[i] *= om;
          return;
        }
      std::vector<T> t1(sz), t2(sz);
      L.vmult(t1.data(), u);
      for (long i = 0; i < sz; ++i)
        t1[i] = rhs[i] - t1[i];
      L.vanka(t2.data(), t1.data());
      for (long i = 0; i < sz; ++i)
        u[i] += om * t2[i];
    }
    int steps_at(int l) const
    {
      const int L = (int)lv.size() - 1;
      return spec.variable ? (1 << (L - l)) : 1;
    }
    // MGSmootherPrecondition::apply
    void smoother_apply(int l, T *u, const T *rhs) const
    {
      const int s2 = steps_at(l);
      for (int i = 0; i < spec.smoothing_steps; ++i)
        relax(l, u, rhs, i == 0);
      for (int k = 1; k < s2; ++k)
        for (int i = 1; i <= spec.smoothing_steps; ++i)
          relax(l, u, rhs, false);
    }
    // MGSmootherPrecondition::smooth
    void smoother_smooth(int l, T *u, const T *rhs) const
    {
      const int s2 = steps_at(l);
      for (int k = 0; k < s2; ++k)
        for (int i = 1; i <= spec.smoothing_steps; ++i)
          relax(l, u, rhs, false);
    }

    // Multigrid::level_v_step
    void level_v_step(int l, std::vector<std::vector<T>> &sol,
                      std::vector<std::vector<T>> &def) const
    {
      if (l == 0)
        {
          smoother_apply(0, sol[0].data(), def[0].data());
          return;
        }
      smoother_apply(l, sol[l].data(), def[l].data());
      const long     sz = lv[l].size();
      std::vector<T> t(sz);
      lv[l].vmult(t.data(), sol[l].data());
      for (long i = 0; i < sz; ++i)
        t[i] = def[l][i] - t[i];
      std::fill(def[l - 1].begin(), def[l - 1].end(), T(0));
      tr[l].restrict_and_add(def[l - 1].data(), t.data());
      std::fill(sol[l - 1].begin(), sol[l - 1].end(), T(0));
      level_v_step(l - 1, sol, def);
      tr[l].prolongate_and_add(sol[l].data(), sol[l - 1].data());
      smoother_smooth(l, sol[l].data(), def[l].data());
    }

    // GMG::vmult (gmg.h:846-859) with FP64 <-> T conversion
    void gmg_vmult(double *dst, const double *src) const
    {
      const int                   nl = (int)lv.size();
      std::vector<std::vector<T>> sol(nl), def(nl);
      for (int l = 0; l < nl; ++l)
        {
          sol[l].assign(lv[l].size(), T(0));
          def[l].assign(lv[l].size(), T(0));
        }
      const long sz = lv[nl - 1].size();
      for (long i = 0; i < sz; ++i)
        def[nl - 1][i] = T(src[i]);
      level_v_step(nl - 1, sol, def);
      for (long i = 0; i < sz; ++i)
        dst[i] = double(sol[nl - 1][i]);
    }

    // SolverFGMRES::solve with ReductionControl (deal.II 9.5 semantics).
    // Returns 0 on success, -1 on no convergence.
    int fgmres(double *x, const double *b)
    {
      const long                       sz = fine.size();
      const int                        m  = spec.basis;
      std::vector<std::vector<double>> v(m + 1, std::vector<double>(sz)),
        z(m + 1, std::vector<double>(sz));
      std::vector<double> aux(sz);
      res_hist.clear();
      int    acc = 0;
      double reduced_tol = 0;
      int    state = 0; // 0 iterate, 1 success, -1 failure
      auto   check = [&](int step, double val) {
        res_hist.push_back(val);
        if (step == 0)
          reduced_tol = val * spec.reltol;
        last_its = step;
        if (val <= reduced_tol)
          return 1;
        if (val <= spec.abstol)
          return 1;
        if (step >= spec.maxiter || std::isnan(val))
          return -1;
        return 0;
      };
      auto dot = [&](const std::vector<double> &a, const std::vector<double> &c) {
        double s = 0;
        for (long i = 0; i < sz; ++i)
          s += a[i] * c[i];
        return s;
      };
      do
        {
          fine.vmult(aux.data(), x);
          for (long i = 0; i < sz; ++i)
            aux[i] = b[i] - aux[i];
          double beta = std::sqrt(dot(aux, aux));
          state       = check(acc, beta);
          if (state != 0)
            break;
          std::vector<std::vector<double>> H(m + 1, std::vector<double>(m, 0.0));
          std::vector<double>              y;
          double                           a = beta;
          for (int j = 0; j < m; ++j)
            {
              for (long i = 0; i < sz; ++i)
                v[j][i] = (a != 0) ? aux[i] / a : 0.0;
              gmg_vmult(z[j].data(), v[j].data());
              fine.vmult(aux.data(), z[j].data());
              // modified Gram-Schmidt (add_and_dot)
              H[0][j] = dot(aux, v[0]);
              for (int i = 1; i <= j; ++i)
                {
                  for (long k = 0; k < sz; ++k)
                    aux[k] -= H[i - 1][j] * v[i - 1][k];
                  H[i][j] = dot(aux, v[i]);
                }
              for (long k = 0; k < sz; ++k)
                aux[k] -= H[j][j] * v[j][k];
              a           = std::sqrt(dot(aux, aux));
              H[j + 1][j] = a;
              int ncol    = (spec.fgmres_variant == 0) ? j : j + 1;
              if (ncol > 0)
                {
                  // least squares min || beta e1 - H(0:ncol, 0:ncol-1) y ||
                  // by Givens rotations
                  std::vector<std::vector<double>> R(ncol + 1, std::vector<double>(ncol));
                  for (int r0 = 0; r0 <= ncol; ++r0)
                    for (int c0 = 0; c0 < ncol; ++c0)
                      R[r0][c0] = H[r0][c0];
                  std::vector<double> g(ncol + 1, 0.0);
                  g[0] = beta;
                  for (int c0 = 0; c0 < ncol; ++c0)
                    {
                      double h1 = R[c0][c0], h2 = R[c0 + 1][c0];
                      double rr = std::hypot(h1, h2);
                      double cs = rr == 0 ? 1 : h1 / rr, sn = rr == 0 ? 0 : h2 / rr;
                      for (int c1 = c0; c1 < ncol; ++c1)
                        {
                          double t1 = R[c0][c1], t2 = R[c0 + 1][c1];
                          R[c0][c1]     = cs * t1 + sn * t2;
                          R[c0 + 1][c1] = -sn * t1 + cs * t2;
                        }
                      double g1 = g[c0], g2 = g[c0 + 1];
                      g[c0]     = cs * g1 + sn * g2;
                      g[c0 + 1] = -sn * g1 + cs * g2;
                    }
                  y.assign(ncol, 0.0);
                  for (int c0 = ncol - 1; c0 >= 0; --c0)
                    {
                      double s = g[c0];
                      for (int c1 = c0 + 1; c1 < ncol; ++c1)
                        s -= R[c0][c1] * y[c1];
                      y[c0] = s / R[c0][c0];
                    }
                  double res = std::abs(g[ncol]);
                  state      = check(++acc, res);
                  if (state != 0)
                    break;
                }
            }
          for (size_t j = 0; j < y.size(); ++j)
            for (long i = 0; i < sz; ++i)
              x[i] += y[j] * z[j][i];
        }
      while (state == 0);
      return state == 1 ? 0 : -1;
    }

    // Gaussian-bump source h(x) (rhs=3, time profile 1): bump b has centre
    // c_b and amplitude a_b from the 64-bit Mersenne twister seeded with
    // `seed`, each variate u = (top 53 bits) / 2^53, c_b = 0.2 + 0.6 u per
    // coordinate (x, y, z drawn in that order even in 2D), a_b = 0.5 + u;
    // h(x) = sum_b a_b exp(-|x - c_b|^2 / 0.02).  The benchmark's synthetic
    // forcing: not a reference source (the reference has the manufactured
    // source of exact_solution.h:53-75 only).
    std::vector<std::array<double, 4>> bumps;
    void make_bumps()
    {
      std::mt19937_64 gen{static_cast<unsigned long long>(spec.seed)};
      bumps.assign(std::max(1, spec.n_bumps), {0, 0, 0, 0});
      for (auto &b : bumps)
        {
          for (int d = 0; d < 4; ++d)
            {
              const double u = std::ldexp(double(gen() >> 11), -53);
              b[d] = d < 3 ? 0.2 + 0.6 * u : 0.5 + u;
            }
        }
    }
    double bump_value(const double *x) const
    {
      double v = 0.0;
      for (const auto &b : bumps)
        {
          double d2 = 0.0;
          for (int d = 0; d < spec.dim; ++d)
            d2 += (x[d] - b[d]) * (x[d] - b[d]);
          v += b[3] * std::exp(-50.0 * d2);
        }
      return v;
    }

    // manufactured heat source, exact_solution.h:53-75 (frequency f)
    double rhs_value(double t, const double *x) const
    {
      if (spec.rhs == 1)
        return 0.0;
      if (spec.rhs == 3)
        return bump_value(x);
      if (spec.problem == 1)
        {
          // wave::RHSFunction (exact_solution.h:166-180)
          const double f = spec.frequency, PI = M_PI;
          double v = std::pow(2.0, spec.dim) * (PI * f) * (PI * f) * std::sin(2 * PI * f * t);
          for (int d = 0; d < spec.dim; ++d)
            v *= std::sin(2 * PI * f * x[d]);
          return v;
        }
      const double f = spec.frequency, PI = M_PI;
      double       v = spec.dim * 4 * PI * PI * f * f * std::sin(2 * PI * f * t) +
                 2 * PI * f * std::cos(2 * PI * f * t);
      for (int d = 0; d < spec.dim; ++d)
        v *= std::sin(2 * PI * f * x[d]);
      return v;
    }
    // VectorTools::interpolate of wave::ExactSolutionV at t (exact_solution.h:
    // 143-163) onto the fine nodes (support points through the Q1 mapping)
    std::vector<double> interpolate_exact_v(double t) const
    {
      const Space        &sp = fine.sp;
      const int           dim = sp.dim, n = sp.n;
      std::vector<double> out(sp.ndofs, 0.0);
      std::vector<long>   idx;
      const double        f = spec.frequency, PI = M_PI;
      for (long c = 0; c < sp.g.n_cells(); ++c)
        {
          sp.cell_dofs(c, idx);
          double V[8][3];
          sp.cell_vertices(c, V);
          int l = 0;
          for (int c3 = 0; c3 < (dim == 3 ? n : 1); ++c3)
            for (int b = 0; b < n; ++b)
              for (int a = 0; a < n; ++a, ++l)
                {
                  double xi[3] = {sh.gll[a], sh.gll[b], dim == 3 ? sh.gll[c3] : 0.0};
                  double x[3], J[3][3];
                  q1_map(dim, V, xi, x, J);
                  double v = 2 * PI * f * std::cos(2 * PI * f * t);
                  for (int d = 0; d < dim; ++d)
                    v *= std::sin(2 * PI * f * x[d]);
                  out[idx[l]] = sp.constrained[idx[l]] ? 0.0 : v;
                }
        }
      return out;
    }

    double exact_value(double t, const double *x) const
    {
      const double f = spec.frequency, PI = M_PI;
      double       v = std::sin(2 * PI * f * t);
      for (int d = 0; d < spec.dim; ++d)
        v *= std::sin(2 * PI * f * x[d]);
      return v;
    }
    void exact_gradient(double t, const double *x, double *g) const
    {
      const double f = spec.frequency, PI = M_PI;
      double       tv = 2 * PI * f * std::sin(2 * PI * f * t);
      for (int i = 0; i < spec.dim; ++i)
        {
          g[i] = tv;
          for (int j = 0; j < spec.dim; ++j)
            g[i] *= (i == j) ? std::cos(2 * PI * f * x[j]) : std::sin(2 * PI * f * x[j]);
        }
    }

    // VectorTools::create_right_hand_side with QGauss(p+1), Dirichlet rows 0.
    void integrate_rhs(double t, double *F) const
    {
      const Space &sp = fine.sp;
      std::fill(F, F + sp.ndofs, 0.0);
      const int         dim = sp.dim, n = sp.n, nq = (int)sh.gq.size(), nl = sp.nc_local;
      std::vector<long> idx;
      for (long c = 0; c < sp.g.n_cells(); ++c)
        {
          sp.cell_dofs(c, idx);
          double V[8][3];
          sp.cell_vertices(c, V);
          for (int qz = 0; qz < (dim == 3 ? nq : 1); ++qz)
            for (int qy = 0; qy < nq; ++qy)
              for (int qx = 0; qx < nq; ++qx)
                {
                  double xi[3] = {sh.gq[qx], sh.gq[qy], dim == 3 ? sh.gq[qz] : 0.0};
                  double w     = sh.gw[qx] * sh.gw[qy] * (dim == 3 ? sh.gw[qz] : 1.0);
                  double x[3], J[3][3], Ji[3][3];
                  q1_map(dim, V, xi, x, J);
                  double JxW = std::abs(det_inv(dim, J, Ji)) * w;
                  double fv  = rhs_value(t, x) * JxW;
                  int    l   = 0;
                  for (int c3 = 0; c3 < (dim == 3 ? n : 1); ++c3)
                    for (int b = 0; b < n; ++b)
                      for (int a = 0; a < n; ++a, ++l)
                        {
                          double phi = sh.S[qx * n + a] * sh.S[qy * n + b] *
                                       (dim == 3 ? sh.S[qz * n + c3] : 1.0);
                          if (!sp.constrained[idx[l]])
                            F[idx[l]] += fv * phi;
                        }
                }
          (void)nl;
        }
    }

    // ErrorCalculator::evaluate_error (exact_solution.h:218-299):
    // QGauss(r+1) in space and time; returns {L2^2, Linf, H1semi^2} summed
    // over the slab with time weights tau*w_q.
    void slab_error(double time, const std::vector<double> &x, const std::vector<double> &prev,
                    double out[3]) const
    {
      const Space &sp = fine.sp;
      const long   N  = sp.ndofs;
      Shape1D      eq;
      eq.init(spec.p, spec.r + 1);
      const auto     tq = gauss_points(spec.r + 1), tw = gauss_weights(spec.r + 1);
      Lagrange       tb = time_basis(spec.type, spec.r);
      const int      ni = (spec.type == DG) ? spec.r + 1 : spec.r;
      const int      dim = sp.dim, n = sp.n, nq = (int)eq.gq.size();
      std::vector<double> u(N);
      std::vector<long>   idx;
      out[0] = 0;
      out[1] = -1;
      out[2] = 0;
      for (int it = 0; it < spec.nts; ++it)
        for (size_t q = 0; q < tq.size(); ++q)
          {
            const double t = time + tau * it + tq[q] * tau;
            std::fill(u.begin(), u.end(), 0.0);
            for (int i = 0; i < tb.size(); ++i)
              {
                double v = tb.value(i, tq[q]);
                if (v == 0.0)
                  continue;
                const double *src;
                if (spec.type == DG)
                  src = x.data() + (long)(it * ni + i) * N;
                else if (it * ni + i == 0)
                  src = prev.data();
                else
                  src = x.data() + (long)(it * ni + i - 1) * N;
                if (spec.type == CGP && i == 0 && it > 0)
                  src = x.data() + (long)(it * ni - 1) * N;
                for (long k = 0; k < N; ++k)
                  u[k] += v * src[k];
              }
            for (long k = 0; k < N; ++k)
              if (sp.constrained[k])
                u[k] = 0.0;
            double l2 = 0, l8 = 0, h1 = 0;
            for (long c = 0; c < sp.g.n_cells(); ++c)
              {
                sp.cell_dofs(c, idx);
                double V[8][3];
                sp.cell_vertices(c, V);
                double cl2 = 0, ch1 = 0;
                for (int qz = 0; qz < (dim == 3 ? nq : 1); ++qz)
                  for (int qy = 0; qy < nq; ++qy)
                    for (int qx = 0; qx < nq; ++qx)
                      {
                        double xi[3] = {eq.gq[qx], eq.gq[qy], dim == 3 ? eq.gq[qz] : 0.0};
                        double w = eq.gw[qx] * eq.gw[qy] * (dim == 3 ? eq.gw[qz] : 1.0);
                        double x0[3], J[3][3], Ji[3][3];
                        q1_map(dim, V, xi, x0, J);
                        double JxW = std::abs(det_inv(dim, J, Ji)) * w;
                        double uh = 0, gh[3] = {0, 0, 0};
                        int    l  = 0;
                        for (int c3 = 0; c3 < (dim == 3 ? n : 1); ++c3)
                          for (int b = 0; b < n; ++b)
                            for (int a = 0; a < n; ++a, ++l)
                              {
                                double sx = eq.S[qx * n + a], sy = eq.S[qy * n + b];
                                double sz = dim == 3 ? eq.S[qz * n + c3] : 1.0;
                                double dx = eq.D[qx * n + a], dy = eq.D[qy * n + b];
                                double dz = dim == 3 ? eq.D[qz * n + c3] : 0.0;
                                double g[3] = {dx * sy * sz, sx * dy * sz, sx * sy * dz};
                                double val  = u[idx[l]];
                                uh += val * sx * sy * sz;
                                for (int d = 0; d < dim; ++d)
                                  {
                                    double s = 0;
                                    for (int e = 0; e < dim; ++e)
                                      s += Ji[e][d] * g[e];
                                    gh[d] += val * s;
                                  }
                              }
                        double ue = exact_value(t, x0), ge[3];
                        exact_gradient(t, x0, ge);
                        double e = uh - ue;
                        cl2 += e * e * JxW;
                        l8 = std::max(l8, std::abs(e));
                        for (int d = 0; d < dim; ++d)
                          ch1 += (gh[d] - ge[d]) * (gh[d] - ge[d]) * JxW;
                      }
                l2 += cl2;
                h1 += ch1;
              }
            out[0] += tau * tw[q] * l2;
            out[1] = std::max(out[1], l8);
            out[2] += tau * tw[q] * h1;
          }
    }

    // TimeIntegratorHeat::solve (time_integrators.h:157-180)
    // TimeIntegratorWave::solve (time_integrators.h:237-280): the
    // displacement slab (heat path with the wave weights plus
    // rhs_matrix_v * prev_v), then the velocities
    //   v_it = AixB u_it + AixG prev_u_   (dG)
    //   v_it = AixB u_it + AixG prev_v_ + AixZ prev_u_   (cG)
    // with prev_u_/prev_v_ the previous step's last block.
    int solve_slab_wave(std::vector<double> &x, std::vector<double> &v,
                        const std::vector<double> &prev, const std::vector<double> &prev_v,
                        double time)
    {
      const int  rc = solve_slab(x, prev, time, &prev_v);
      const long N  = fine.N();
      const int  ni = AixB.m;
      std::fill(v.begin(), v.end(), 0.0);
      for (int it = 0; it < spec.nts; ++it)
        {
          const double *pu = (it == 0) ? prev.data() : x.data() + (long)(it * ni - 1) * N;
          const double *pv = (it == 0) ? prev_v.data() : v.data() + (long)(it * ni - 1) * N;
          for (int i = 0; i < ni; ++i)
            {
              double *vi = v.data() + (long)(it * ni + i) * N;
              for (int j = 0; j < ni; ++j)
                if (AixB(i, j) != 0.0)
                  for (long k = 0; k < N; ++k)
                    vi[k] += AixB(i, j) * x[(long)(it * ni + j) * N + k];
              if (spec.type == DG)
                {
                  if (AixG(i, 0) != 0.0)
                    for (long k = 0; k < N; ++k)
                      vi[k] += AixG(i, 0) * pu[k];
                }
              else
                {
                  if (AixG(i, 0) != 0.0)
                    for (long k = 0; k < N; ++k)
                      vi[k] += AixG(i, 0) * pv[k];
                  if (AixZ(i, 0) != 0.0)
                    for (long k = 0; k < N; ++k)
                      vi[k] += AixZ(i, 0) * pu[k];
                }
            }
        }
      return rc;
    }

    int solve_slab(std::vector<double> &x, const std::vector<double> &prev, double time,
                   const std::vector<double> *prev_v = nullptr)
    {
      const long          N = fine.N();
      std::vector<double> rhs(fine.size(), 0.0), tmp(N);
      // wave: rhs_matrix_v.vmult_add(rhs, prev_v), mass part only
      if (prev_v && !rhsV.all_zero())
        {
          spmv_free(fine.M, fine.Mv, fine.sp.constrained, prev_v->data(), tmp.data());
          for (int j = 0; j < nt; ++j)
            if (rhsV(j, 0) != 0.0)
              for (long k = 0; k < N; ++k)
                rhs[j * N + k] += rhsV(j, 0) * tmp[k];
        }
      // rhs_matrix.vmult(rhs, prev_x) (operators.h:154-187)
      if (!rhsK.all_zero())
        {
          spmv_free(fine.K, fine.Kv, fine.sp.constrained, prev.data(), tmp.data());
          for (int j = 0; j < nt; ++j)
            if (rhsK(j, 0) != 0.0)
              for (long k = 0; k < N; ++k)
                rhs[j * N + k] += rhsK(j, 0) * tmp[k];
        }
      if (!rhsM.all_zero())
        {
          spmv_free(fine.M, fine.Mv, fine.sp.constrained, prev.data(), tmp.data());
          for (int j = 0; j < nt; ++j)
            if (rhsM(j, 0) != 0.0)
              for (long k = 0; k < N; ++k)
                rhs[j * N + k] += rhsM(j, 0) * tmp[k];
        }
      // assemble_force (time_integrators.h:60-87)
      if (spec.rhs != 1)
        {
          Lagrange   tb = time_basis(spec.type, spec.r);
          const int  ni = Alpha1.m;
          for (int it = 0; it < spec.nts; ++it)
            for (int j = 0; j < tb.size(); ++j)
              {
                double t = time + tau * it + tau * tb.x[j];
                integrate_rhs(t, tmp.data());
                if (spec.type == DG)
                  {
                    for (long k = 0; k < N; ++k)
                      rhs[(long)(j + it * ni) * N + k] += Alpha1(j, j) * tmp[k];
                  }
                else if (j == 0)
                  {
                    for (int i = 0; i < Gamma1.m; ++i)
                      for (long k = 0; k < N; ++k)
                        rhs[(long)(i + it * ni) * N + k] += -Gamma1(i, 0) * tmp[k];
                  }
                else
                  for (long k = 0; k < N; ++k)
                    rhs[(long)(j - 1 + it * ni) * N + k] += Alpha1(j - 1, j - 1) * tmp[k];
              }
        }
      for (int j = 0; j < nt; ++j)
        for (long k = 0; k < N; ++k)
          x[j * N + k] = spec.extrapolate ? prev[k] : 0.0;
      int rc = fgmres(x.data(), rhs.data());
      for (int j = 0; j < nt; ++j)
        for (long k = 0; k < N; ++k)
          if (fine.sp.constrained[k])
            x[j * N + k] = 0.0;
      return rc;
    }
  };

  struct AnyProblem
  {
    bool                             fp32 = false;
    std::unique_ptr<Problem<double>> pd;
    std::unique_ptr<Problem<float>>  pf;
  };
} // namespace orc

using namespace orc;

  template <typename PT>
static void lvl_apply(PT &P, int l, int what, double *dst, const double *src)
{
  using T  = typename std::remove_reference<decltype(P.lv[0].At[0])>::type;
  auto &L  = P.lv[l];
  long  sz = L.size();
  if (what == 0 || what == 1)
    {
      std::vector<T> s(src, src + sz), d(sz);
      if (what == 0)
        L.vmult(d.data(), s.data());
      else
        L.vanka(d.data(), s.data());
      for (long i = 0; i < sz; ++i)
        dst[i] = d[i];
    }
  else
    {
      // 2: prolongate_and_add from level l-1 into l; 3: restrict_and_add l -> l-1
      long           szc = P.lv[l - 1].size();
      const auto    &t   = P.tr[l];
      if (what == 2)
        {
          std::vector<T> s(src, src + szc), d(dst, dst + sz);
          t.prolongate_and_add(d.data(), s.data());
          for (long i = 0; i < sz; ++i)
            dst[i] = d[i];
        }
      else
        {
          std::vector<T> s(src, src + sz), d(dst, dst + szc);
          t.restrict_and_add(d.data(), s.data());
          for (long i = 0; i < szc; ++i)
            dst[i] = d[i];
        }
    }
}


// ======================================================================= C API
extern "C"
{
  // --- time algebra (all matrices row-major) ---
  int oracle_time_nt(int type, int r, int nts) { return (type == CGP ? r : r + 1) * nts; }

  void oracle_fe_time_weights(int type, int r, double tau, int nts, double *A, double *B,
                              double *G, double *Z)
  {
    auto w = fe_time_weights(type, r, tau, nts);
    std::copy(w[0].a.begin(), w[0].a.end(), A);
    std::copy(w[1].a.begin(), w[1].a.end(), B);
    std::copy(w[2].a.begin(), w[2].a.end(), G);
    std::copy(w[3].a.begin(), w[3].a.end(), Z);
  }
  void oracle_dg_weights(int r, double *mass, double *der, double *jump)
  {
    auto w = dg_weights(r);
    std::copy(w[0].a.begin(), w[0].a.end(), mass);
    std::copy(w[1].a.begin(), w[1].a.end(), der);
    std::copy(w[2].a.begin(), w[2].a.end(), jump);
  }
  void oracle_cg_weights(int r, double *m, double *d)
  {
    auto w = cg_weights(r);
    std::copy(w[0].a.begin(), w[0].a.end(), m);
    std::copy(w[1].a.begin(), w[1].a.end(), d);
  }
  // out sizes: r0,r1: (ni*nts)^2, r2..r4: ni*nts
  void oracle_fe_time_weights_wave(int type, int r, int nts, double *r0, double *r1, double *r2,
                                   double *r3, double *r4)
  {
    auto w  = fe_time_weights(type, r, 1.0, 1);
    auto ww = fe_time_weights_wave(type, w[0], w[1], w[2], w[3], nts);
    double *o[5] = {r0, r1, r2, r3, r4};
    for (int k = 0; k < 5; ++k)
      std::copy(ww[k].a.begin(), ww[k].a.end(), o[k]);
  }
  // tp_02.cc:57-77 variant: wave weights from the raw DG weights
  void oracle_wave_from(int type, int m, const double *A, const double *B, const double *G,
                        const double *Z, int nts, double *r0, double *r1, double *r2, double *r3,
                        double *r4)
  {
    Mat a(m, m), b(m, m), g(m, 1), z(m, 1);
    std::copy(A, A + m * m, a.a.begin());
    std::copy(B, B + m * m, b.a.begin());
    std::copy(G, G + m, g.a.begin());
    if (Z)
      std::copy(Z, Z + m, z.a.begin());
    auto    ww   = fe_time_weights_wave(type, a, b, g, z, nts);
    double *o[5] = {r0, r1, r2, r3, r4};
    for (int k = 0; k < 5; ++k)
      std::copy(ww[k].a.begin(), ww[k].a.end(), o[k]);
  }
  void oracle_time_prolongation(int type, int r, int nts, double *P)
  {
    auto m = time_prolongation(type, r, nts);
    std::copy(m.a.begin(), m.a.end(), P);
  }
  void oracle_time_restriction(int type, int r, int nts, double *R)
  {
    auto m = time_restriction(type, r, nts);
    std::copy(m.a.begin(), m.a.end(), R);
  }
  // returns rows*cols written; rows/cols through pointers
  void oracle_time_projection(int type, int r_src, int r_dst, int nts, double *P, int *rows,
                              int *cols)
  {
    auto m = time_projection(type, r_src, r_dst, nts);
    *rows  = m.m;
    *cols  = m.n;
    if (P)
      std::copy(m.a.begin(), m.a.end(), P);
  }
  void oracle_quadrature(int kind, int n, double *pts, double *wts)
  {
    std::vector<double> p;
    if (kind == 0)
      p = gauss_points(n);
    else if (kind == 1)
      p = lobatto_points(n);
    else
      p = radau_right_points(n);
    Lagrange L;
    L.x = p;
    for (int i = 0; i < n; ++i)
      {
        pts[i] = p[i];
        if (wts)
          wts[i] = integrate([&](double t) { return L.value(i, t); });
      }
  }
  int oracle_mg_sequence(int n_sp, int kmax, int kmin, int nts, int nts_min, int tbs, int *out)
  {
    auto v = time_mg_sequence(n_sp, kmax, kmin, nts, nts_min, tbs != 0);
    if (out)
      std::copy(v.begin(), v.end(), out);
    return (int)v.size();
  }

  // --- problem (hierarchy + solver) ---
  void *oracle_problem_create(const char *spec, const double *coords, const double *kappa)
  {
    Spec s  = parse_spec(spec);
    auto *h = new AnyProblem;
    h->fp32 = s.mg_float != 0;
    if (h->fp32)
      {
        h->pf = std::make_unique<Problem<float>>();
        h->pf->init(s, coords, kappa);
      }
    else
      {
        h->pd = std::make_unique<Problem<double>>();
        h->pd->init(s, coords, kappa);
      }
    return h;
  }
  void oracle_problem_destroy(void *h) { delete (AnyProblem *)h; }

#define ORC_DISPATCH(h, ...)                                                                     \
  do                                                                                              \
    {                                                                                             \
      auto *ap_ = (AnyProblem *)(h);                                                              \
      if (ap_->fp32)                                                                              \
        {                                                                                         \
          auto &P = *ap_->pf;                                                                     \
          __VA_ARGS__;                                                                            \
        }                                                                                         \
      else                                                                                        \
        {                                                                                         \
          auto &P = *ap_->pd;                                                                     \
          __VA_ARGS__;                                                                            \
        }                                                                                         \
    }                                                                                             \
  while (0)

  // info: [n_levels, nt_fine, ndofs_fine, n_cells_fine, tau*1e? -> use oracle_tau]
  int oracle_n_levels(void *h)
  {
    int r = 0;
    ORC_DISPATCH(h, r = (int)P.lv.size());
    return r;
  }
  double oracle_tau(void *h)
  {
    double r = 0;
    ORC_DISPATCH(h, r = P.tau);
    return r;
  }
  // per level: [ndofs, nt, ncells, r, nts, mg_type(to this level), nc0, nc1, nc2]
  void oracle_level_info(void *h, int l, long *info)
  {
    ORC_DISPATCH(h, {
      auto &L = P.lv[l];
      info[0] = L.N();
      info[1] = L.nt;
      info[2] = L.sp.g.n_cells();
      info[3] = L.r;
      info[4] = L.nts;
      info[5] = l > 0 ? P.mg_type[l - 1] : 0;
      info[6] = L.sp.g.nc[0];
      info[7] = L.sp.g.nc[1];
      info[8] = L.sp.g.nc[2];
    });
  }
  double oracle_level_omega(void *h, int l)
  {
    double r = 0;
    ORC_DISPATCH(h, r = P.lv[l].omega);
    return r;
  }
  void oracle_level_time_matrices(void *h, int l, double *A, double *B)
  {
    ORC_DISPATCH(h, {
      auto &L = P.lv[l];
      std::copy(L.Alpha.a.begin(), L.Alpha.a.end(), A);
      std::copy(L.Beta.a.begin(), L.Beta.a.end(), B);
    });
  }
  // FP64 fine system vmult / Tvmult, block-major
  void oracle_fine_vmult(void *h, double *dst, const double *src, int transpose)
  {
    ORC_DISPATCH(h, P.fine.vmult(dst, src, transpose != 0));
  }
  // level operator / Vanka / transfers in double (inputs converted)
  void oracle_level_apply(void *h, int l, int what, double *dst, const double *src)
  {
    ORC_DISPATCH(h, lvl_apply(P, l, what, dst, src));
  }
  void oracle_gmg_vmult(void *h, double *dst, const double *src)
  {
    ORC_DISPATCH(h, P.gmg_vmult(dst, src));
  }
  // assembled fine K or M (CSR) export: returns nnz; arrays may be null
  // threads the oracle's parallel loops use (OpenMP)
  int oracle_num_threads(void)
  {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
  }
  // the load vector int f(t, x) phi_i dx of the fine level (Dirichlet
  // entries zero), the spatial part of assemble_force (time_integrators.h:60-87)
  void oracle_load_vector(void *h, double time, double *out)
  {
    ORC_DISPATCH(h, P.integrate_rhs(time, out));
  }
  // raw power-iteration estimate of lambda_max(P^-1 A) on level l (0 where
  // no estimate was made)
  double oracle_level_lambda(void *h, int l)
  {
    double out = 0.0;
    ORC_DISPATCH(h, out = P.lambda_log.at(l));
    return out;
  }
  // assembled K (which=0) or M (1) of level l in CSR, FP64 values
  long oracle_level_matrix(void *h, int l, int which, long *rp, long *ci, double *v)
  {
    long nnz = 0;
    ORC_DISPATCH(h, {
      const CSR &A = which == 0 ? P.lv.at(l).K : P.lv.at(l).M;
      nnz          = (long)A.ci.size();
      if (rp)
        for (size_t i = 0; i < A.rp.size(); ++i)
          rp[i] = A.rp[i];
      if (ci)
        for (size_t i = 0; i < A.ci.size(); ++i)
          ci[i] = A.ci[i];
      if (v)
        for (size_t i = 0; i < A.v.size(); ++i)
          v[i] = A.v[i];
    });
    return nnz;
  }
  long oracle_fine_matrix(void *h, int which, long *rp, long *ci, double *v)
  {
    long nnz = 0;
    ORC_DISPATCH(h, {
      const CSR &A = which == 0 ? P.fine.K : P.fine.M;
      nnz          = (long)A.ci.size();
      if (rp)
        std::copy(A.rp.begin(), A.rp.end(), rp);
      if (ci)
        std::copy(A.ci.begin(), A.ci.end(), ci);
      if (v)
        std::copy(A.v.begin(), A.v.end(), v);
    });
    return nnz;
  }
  // One FGMRES solve of A x = b (x holds the initial guess on entry).
  int oracle_fgmres(void *h, double *x, const double *b, int *its, double *hist, int hist_cap)
  {
    int rc = 0;
    ORC_DISPATCH(h, {
      rc   = P.fgmres(x, b);
      *its = P.last_its;
      for (int i = 0; i < (int)P.res_hist.size() && i < hist_cap; ++i)
        hist[i] = P.res_hist[i];
    });
    return rc;
  }
  // One heat slab (time_integrators.h:157-180): x (nt blocks) overwritten.
  int oracle_solve_slab(void *h, double *x, const double *prev, double time, int *its)
  {
    int rc = 0;
    ORC_DISPATCH(h, {
      const long          N = P.fine.N();
      std::vector<double> xv(x, x + (long)P.nt * N), pv(prev, prev + N);
      rc = P.solve_slab(xv, pv, time);
      std::copy(xv.begin(), xv.end(), x);
      *its = P.last_its;
    });
    return rc;
  }
  // One wave slab (time_integrators.h:237-280): x (displacement) and v
  // (velocity), nt blocks each, overwritten.
  int oracle_solve_slab_wave(void *h, double *x, double *v, const double *prev,
                             const double *prev_v, double time, int *its)
  {
    int rc = 0;
    ORC_DISPATCH(h, {
      const long          N = P.fine.N();
      std::vector<double> xv(x, x + (long)P.nt * N), vv((long)P.nt * N), pv(prev, prev + N),
        pvv(prev_v, prev_v + N);
      rc = P.solve_slab_wave(xv, vv, pv, pvv, time);
      std::copy(xv.begin(), xv.end(), x);
      std::copy(vv.begin(), vv.end(), v);
      *its = P.last_its;
    });
    return rc;
  }
  // ErrorCalculator::evaluate_error of one slab (exact_solution.h:218-299):
  // out = {L2^2, Linf, H1semi^2}; x: nt blocks, prev: one block.
  int oracle_slab_error(void *h, const double *x, const double *prev, double time, double *out)
  {
    ORC_DISPATCH(h, {
      const long          N = P.fine.N();
      std::vector<double> xv(x, x + (long)P.nt * N), pv(prev, prev + N);
      P.slab_error(time, xv, pv, out);
    });
    return 0;
  }
  // ---- Discrete adjoint of the heat slab sequence (FP64, dG), the check of
  // the GPU adjoint.  Slab n: A x_n = R E x_{n-1} + f_n from x_{-1} = 0
  // (TimeIntegratorHeat::solve, time_integrators.h:157-180); misfit
  // J = 1/2 sum_n e_n^T (Alpha (x) M) e_n, e_n = x_n - d_n; adjoint
  // A^T lambda_n = (Alpha (x) M) e_n + E^T R^T lambda_{n+1} solved with the
  // problem hT (spec transpose_time=1: FGMRES with the multigrid of the
  // transposed system); dJ/dlog(kappa_c) = -kappa_c sum_n sum_j
  // lambda_{n,j}^T K_c (Alpha x_n)_j with the unit-coefficient element
  // stiffness K_c.  Vectors block-major (block j = time dof j).
  static double misfit_slab(Problem<double> &P, const double *x, const double *d,
                            std::vector<double> &qe)
  {
    const long N = P.fine.N();
    const int  nt = P.nt;
    std::vector<double> e(size_t(N) * nt), tmp(N);
    for (size_t k = 0; k < e.size(); ++k)
      e[k] = x[k] - d[k];
    qe.assign(size_t(N) * nt, 0.0);
    const Mat &A = P.fine.Alpha;
    for (int i = 0; i < nt; ++i)
      {
        spmv_free(P.fine.M, P.fine.Mv, P.fine.sp.constrained, e.data() + size_t(i) * N, tmp.data());
        for (int j = 0; j < nt; ++j)
          if (A(j, i) != 0.0)
            for (long k = 0; k < N; ++k)
              qe[size_t(j) * N + k] += A(j, i) * tmp[k];
      }
    double J = 0.0;
    for (size_t k = 0; k < e.size(); ++k)
      J += e[k] * qe[k];
    return 0.5 * J;
  }
  int oracle_forward_trajectory(void *h, int S, double *traj)
  {
    auto *ap = (AnyProblem *)h;
    if (ap->fp32)
      return -1;
    Problem<double> &P = *ap->pd;
    const long       N = P.fine.N();
    const int        nt = P.nt;
    std::vector<double> x(size_t(N) * nt, 0.0), prev(N);
    for (int n = 0; n < S; ++n)
      {
        std::copy(x.begin() + size_t(nt - 1) * N, x.begin() + size_t(nt) * N, prev.begin());
        if (P.solve_slab(x, prev, n * P.spec.nts * P.tau) != 0)
          return -3;
        std::copy(x.begin(), x.end(), traj + size_t(n) * N * nt);
      }
    return 0;
  }
  int oracle_adjoint_gradient(void *h, void *hT, const double *obs, int S, double *grad,
                              double *misfit)
  {
    auto *ap = (AnyProblem *)h, *at = (AnyProblem *)hT;
    if (ap->fp32 || at->fp32)
      return -1;
    Problem<double> &P = *ap->pd, &PT = *at->pd;
    if (P.spec.type != DG || P.spec.problem != 0 || !PT.spec.transpose_time)
      return -2;
    const long N = P.fine.N();
    const int  nt = P.nt;
    const Space &sp = P.fine.sp;
    std::vector<std::vector<double>> ck(S, std::vector<double>(N));
    std::vector<double> x(size_t(N) * nt, 0.0), qe, lam(size_t(N) * nt), lr(N, 0.0), tmp(N);
    double J = 0.0;
    for (int n = 0; n < S; ++n)
      {
        std::copy(x.begin() + size_t(nt - 1) * N, x.begin() + size_t(nt) * N, ck[n].begin());
        if (P.solve_slab(x, ck[n], n * P.spec.nts * P.tau) != 0)
          return -3;
        J += misfit_slab(P, x.data(), obs + size_t(n) * N * nt, qe);
      }
    const long nc = sp.g.n_cells();
    std::fill(grad, grad + nc, 0.0);
    Coefficient unit; // constant 1
    unit.dim = sp.dim;
    std::vector<long>   idx;
    std::vector<double> Me, Ke;
    const Mat          &A = P.fine.Alpha;
    for (int n = S - 1; n >= 0; --n)
      {
        if (P.solve_slab(x, ck[n], n * P.spec.nts * P.tau) != 0)
          return -3;
        misfit_slab(P, x.data(), obs + size_t(n) * N * nt, qe);
        if (n + 1 < S)
          for (long k = 0; k < N; ++k)
            qe[size_t(nt - 1) * N + k] += lr[k];
        std::fill(lam.begin(), lam.end(), 0.0);
        if (PT.fgmres(lam.data(), qe.data()) != 0)
          return -4;
        for (int j = 0; j < nt; ++j)
          for (long k = 0; k < N; ++k)
            if (sp.constrained[k])
              lam[size_t(j) * N + k] = 0.0;
        // w_j = sum_i Alpha(j,i) x_i
        std::vector<double> w(size_t(N) * nt, 0.0);
        for (int j = 0; j < nt; ++j)
          for (int i = 0; i < nt; ++i)
            if (A(j, i) != 0.0)
              for (long k = 0; k < N; ++k)
                w[size_t(j) * N + k] += A(j, i) * x[size_t(i) * N + k];
        for (long c = 0; c < nc; ++c)
          {
            element_matrices(sp, P.sh, unit, 0, c, Me, Ke);
            sp.cell_dofs(c, idx);
            const int nl = (int)idx.size();
            double    s  = 0.0;
            for (int j = 0; j < nt; ++j)
              for (int a = 0; a < nl; ++a)
                {
                  if (sp.constrained[idx[a]])
                    continue;
                  double r = 0.0;
                  for (int b = 0; b < nl; ++b)
                    if (!sp.constrained[idx[b]])
                      r += Ke[size_t(a) * nl + b] * w[size_t(j) * N + idx[b]];
                  s += lam[size_t(j) * N + idx[a]] * r;
                }
            const double kap = P.coef.kappa.empty() ? 1.0 : P.coef.kappa[c];
            grad[c] -= kap * s;
          }
        // R^T lambda_n = sum_j Gamma_j M lambda_{n,j} (dG: R = Gamma (x) M)
        std::fill(lr.begin(), lr.end(), 0.0);
        for (int j = 0; j < nt; ++j)
          if (P.rhsM(j, 0) != 0.0)
            {
              spmv_free(P.fine.M, P.fine.Mv, sp.constrained, lam.data() + size_t(j) * N, tmp.data());
              for (long k = 0; k < N; ++k)
                lr[k] += P.rhsM(j, 0) * tmp[k];
            }
      }
    *misfit = J;
    return 0;
  }
  int oracle_misfit(void *h, const double *obs, int S, double *misfit)
  {
    auto *ap = (AnyProblem *)h;
    if (ap->fp32)
      return -1;
    Problem<double> &P = *ap->pd;
    const long       N = P.fine.N();
    const int        nt = P.nt;
    std::vector<double> x(size_t(N) * nt, 0.0), prev(N), qe;
    double J = 0.0;
    for (int n = 0; n < S; ++n)
      {
        std::copy(x.begin() + size_t(nt - 1) * N, x.begin() + size_t(nt) * N, prev.begin());
        if (P.solve_slab(x, prev, n * P.spec.nts * P.tau) != 0)
          return -3;
        J += misfit_slab(P, x.data(), obs + size_t(n) * N * nt, qe);
      }
    *misfit = J;
    return 0;
  }
  // Full heat run (tp_01.cc:630-683): stats = [slabs, total_its, L2, Linf,
  // H1semi, n_dofs, n_blocks]; its_per_slab optional (cap entries).
  int oracle_run_heat(void *h, double *stats, int *its_per_slab, int cap)
  {
    int rc = 0;
    ORC_DISPATCH(h, {
      const long          N  = P.fine.N();
      const int           nt = P.nt;
      const bool          wave = P.spec.problem == 1;
      std::vector<double> x(size_t(N) * nt, 0.0), prev(N, 0.0);
      std::vector<double> v(wave ? size_t(N) * nt : 0, 0.0), prev_v(wave ? N : 0, 0.0);
      // initial values (tp_01.cc:566-568): u(0) = 0 for the manufactured
      // solution; the wave velocity interpolates v(0)
      if (wave)
        {
          auto v0 = P.interpolate_exact_v(0.0);
          std::copy(v0.begin(), v0.end(), v.begin() + (long)(nt - 1) * N);
        }
      double time = 0.0;
      int    slab = 0, total = 0;
      double l2 = 0, l8 = -1, h1 = 0;
      while (time < P.spec.end_time)
        {
          std::copy(x.begin() + (long)(nt - 1) * N, x.begin() + (long)nt * N, prev.begin());
          int r;
          if (wave)
            {
              std::copy(v.begin() + (long)(nt - 1) * N, v.begin() + (long)nt * N,
                        prev_v.begin());
              r = P.solve_slab_wave(x, v, prev, prev_v, time);
            }
          else
            r = P.solve_slab(x, prev, time);
          if (r != 0)
            rc = r;
          if (its_per_slab && slab < cap)
            its_per_slab[slab] = P.last_its;
          total += P.last_its;
          double e[3];
          P.slab_error(time, x, prev, e);
          l2 += e[0];
          l8 = std::max(l8, e[1]);
          h1 += e[2];
          time += P.spec.nts * P.tau;
          ++slab;
        }
      stats[0] = slab;
      stats[1] = total;
      stats[2] = std::sqrt(l2);
      stats[3] = l8;
      stats[4] = std::sqrt(h1);
      stats[5] = (double)N;
      stats[6] = nt;
    });
    return rc;
  }
}
