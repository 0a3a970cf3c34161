// Time finite element algebra (see time_fe.h for the reference map).
#include "time_fe.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <stdexcept>

namespace stfem
{
  bool DMat::all_zero() const
  {
    for (double v : a)
      if (v != 0.0)
        return false;
    return true;
  }
  DMat DMat::transposed() const
  {
    DMat t(n, m);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < n; ++j)
        t(j, i) = (*this)(i, j);
    return t;
  }
  DMat DMat::operator*(const DMat &B) const
  {
    if (n != B.m)
      throw std::runtime_error("DMat: shape mismatch");
    DMat C(m, B.n);
    for (int i = 0; i < m; ++i)
      for (int k = 0; k < n; ++k)
        {
          const double aik = (*this)(i, k);
          if (aik != 0.0)
            for (int j = 0; j < B.n; ++j)
              C(i, j) += aik * B(k, j);
        }
    return C;
  }
  DMat &DMat::scale(double s)
  {
    for (double &v : a)
      v *= s;
    return *this;
  }
  // LU with partial pivoting, then solve for the identity.
  DMat DMat::inverse() const
  {
    if (m != n)
      throw std::runtime_error("DMat::inverse: not square");
    std::vector<double> lu = a;
    std::vector<int>    piv(n);
    for (int k = 0; k < n; ++k)
      {
        int    p = k;
        double best = std::abs(lu[size_t(k) * n + k]);
        for (int i = k + 1; i < n; ++i)
          if (std::abs(lu[size_t(i) * n + k]) > best)
            best = std::abs(lu[size_t(i) * n + k]), p = i;
        if (best == 0.0)
          throw std::runtime_error("DMat::inverse: singular");
        piv[k] = p;
        if (p != k)
          for (int j = 0; j < n; ++j)
            std::swap(lu[size_t(k) * n + j], lu[size_t(p) * n + j]);
        for (int i = k + 1; i < n; ++i)
          {
            double f = lu[size_t(i) * n + k] /= lu[size_t(k) * n + k];
            for (int j = k + 1; j < n; ++j)
              lu[size_t(i) * n + j] -= f * lu[size_t(k) * n + j];
          }
      }
    DMat inv(n, n);
    for (int c = 0; c < n; ++c)
      {
        std::vector<double> b(n, 0.0);
        b[c] = 1.0;
        for (int k = 0; k < n; ++k)
          if (piv[k] != k)
            std::swap(b[k], b[piv[k]]);
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < i; ++j)
            b[i] -= lu[size_t(i) * n + j] * b[j];
        for (int i = n - 1; i >= 0; --i)
          {
            for (int j = i + 1; j < n; ++j)
              b[i] -= lu[size_t(i) * n + j] * b[j];
            b[i] /= lu[size_t(i) * n + i];
          }
        for (int i = 0; i < n; ++i)
          inv(i, c) = b[i];
      }
    return inv;
  }

  namespace
  {
    // Cyclic Jacobi eigen-solver for a small dense symmetric matrix.
    // On return `A` is overwritten; eigenvalues in `lam`, eigenvectors as the
    // columns of V (row-major n x n).
    void sym_eig(std::vector<double> A, int n, std::vector<double> &lam, std::vector<double> &V)
    {
      V.assign(size_t(n) * n, 0.0);
      for (int i = 0; i < n; ++i)
        V[size_t(i) * n + i] = 1.0;
      for (int sweep = 0; sweep < 100; ++sweep)
        {
          double off = 0.0;
          for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j)
              off += A[size_t(i) * n + j] * A[size_t(i) * n + j];
          if (off < 1e-32)
            break;
          for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q)
              {
                const double apq = A[size_t(p) * n + q];
                if (std::abs(apq) < 1e-300)
                  continue;
                const double app = A[size_t(p) * n + p], aqq = A[size_t(q) * n + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) /
                                 (std::abs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k)
                  {
                    const double akp = A[size_t(k) * n + p], akq = A[size_t(k) * n + q];
                    A[size_t(k) * n + p] = c * akp - s * akq;
                    A[size_t(k) * n + q] = s * akp + c * akq;
                  }
                for (int k = 0; k < n; ++k)
                  {
                    const double apk = A[size_t(p) * n + k], aqk = A[size_t(q) * n + k];
                    A[size_t(p) * n + k] = c * apk - s * aqk;
                    A[size_t(q) * n + k] = s * apk + c * aqk;
                  }
                for (int k = 0; k < n; ++k)
                  {
                    const double vkp = V[size_t(k) * n + p], vkq = V[size_t(k) * n + q];
                    V[size_t(k) * n + p] = c * vkp - s * vkq;
                    V[size_t(k) * n + q] = s * vkp + c * vkq;
                  }
              }
        }
      lam.resize(n);
      for (int i = 0; i < n; ++i)
        lam[i] = A[size_t(i) * n + i];
    }

    // Zeros of the Jacobi polynomial P_m^{(a,b)} on [-1,1] from the
    // eigenvalues of its symmetric Jacobi (recurrence) matrix; also the
    // Gauss-Jacobi weights (first eigenvector components) times mu0.
    void gauss_jacobi(int m, double a, double b, std::vector<double> &x, std::vector<double> &w)
    {
      x.clear();
      w.clear();
      if (m <= 0)
        return;
      std::vector<double> J(size_t(m) * m, 0.0);
      for (int k = 0; k < m; ++k)
        {
          const double s = 2.0 * k + a + b;
          double       diag;
          if (k == 0)
            diag = (b - a) / (a + b + 2.0);
          else
            diag = (b * b - a * a) / (s * (s + 2.0));
          J[size_t(k) * m + k] = diag;
          if (k + 1 < m)
            {
              const double kk = k + 1.0, s1 = 2.0 * kk + a + b;
              const double num = 4.0 * kk * (kk + a) * (kk + b) * (kk + a + b);
              const double den = s1 * s1 * (s1 + 1.0) * (s1 - 1.0);
              const double off = std::sqrt(num / den);
              J[size_t(k) * m + k + 1] = off;
              J[size_t(k + 1) * m + k] = off;
            }
        }
      std::vector<double> lam, V;
      sym_eig(J, m, lam, V);
      const double mu0 = std::pow(2.0, a + b + 1.0) * std::tgamma(a + 1.0) * std::tgamma(b + 1.0) /
                         std::tgamma(a + b + 2.0);
      std::vector<int> ord(m);
      for (int i = 0; i < m; ++i)
        ord[i] = i;
      std::sort(ord.begin(), ord.end(), [&](int i, int j) { return lam[i] < lam[j]; });
      for (int i : ord)
        {
          x.push_back(lam[i]);
          w.push_back(mu0 * V[i] * V[i]);
        }
    }

    // weights of an interpolatory rule on [0,1] by integrating its Lagrange basis
    std::vector<double> interpolatory_weights(const std::vector<double> &nodes)
    {
      const Quad1D  g = quad_gauss(int(nodes.size()) + 2);
      LagrangeBasis L{nodes};
      std::vector<double> w(nodes.size(), 0.0);
      for (size_t i = 0; i < nodes.size(); ++i)
        for (size_t q = 0; q < g.x.size(); ++q)
          w[i] += g.w[q] * L.value(int(i), g.x[q]);
      return w;
    }
  } // namespace

  Quad1D quad_gauss(int n)
  {
    std::vector<double> x, w;
    gauss_jacobi(n, 0.0, 0.0, x, w);
    Quad1D q;
    for (int i = 0; i < n; ++i)
      {
        q.x.push_back(0.5 * (x[i] + 1.0));
        q.w.push_back(0.5 * w[i]);
      }
    return q;
  }

  // interior Lobatto nodes are the zeros of P_{n-2}^{(1,1)}
  Quad1D quad_gauss_lobatto(int n)
  {
    if (n < 2)
      throw std::runtime_error("Gauss-Lobatto needs n >= 2");
    std::vector<double> x, w;
    gauss_jacobi(n - 2, 1.0, 1.0, x, w);
    Quad1D q;
    q.x.push_back(0.0);
    for (double v : x)
      q.x.push_back(0.5 * (v + 1.0));
    q.x.push_back(1.0);
    q.w = interpolatory_weights(q.x);
    return q;
  }

  // Radau nodes including the right end point: the remaining n-1 nodes are
  // the zeros of P_{n-1}^{(1,0)}
  Quad1D quad_gauss_radau_right(int n)
  {
    std::vector<double> x, w;
    gauss_jacobi(n - 1, 1.0, 0.0, x, w);
    Quad1D q;
    for (double v : x)
      q.x.push_back(0.5 * (v + 1.0));
    q.x.push_back(1.0);
    q.w = interpolatory_weights(q.x);
    return q;
  }

  double LagrangeBasis::value(int i, double t) const
  {
    double v = 1.0;
    for (int k = 0; k < size(); ++k)
      if (k != i)
        v *= (t - nodes[k]) / (nodes[i] - nodes[k]);
    return v;
  }
  double LagrangeBasis::deriv(int i, double t) const
  {
    // product rule, sum over the dropped factor
    double s = 0.0;
    for (int k = 0; k < size(); ++k)
      {
        if (k == i)
          continue;
        double v = 1.0 / (nodes[i] - nodes[k]);
        for (int m = 0; m < size(); ++m)
          if (m != i && m != k)
            v *= (t - nodes[m]) / (nodes[i] - nodes[m]);
        s += v;
      }
    return s;
  }

  LagrangeBasis get_time_basis(TimeStepType type, int r)
  {
    LagrangeBasis L;
    L.nodes = (type == TimeStepType::CGP) ? quad_gauss_lobatto(r + 1).x :
                                            quad_gauss_radau_right(r + 1).x;
    return L;
  }

  DGWeights get_dg_weights(int r)
  {
    const LagrangeBasis L = get_time_basis(TimeStepType::DG, r);
    const Quad1D        q = quad_gauss(r + 2);
    const int           n = r + 1;
    DGWeights           W{DMat(n, n), DMat(n, n), DMat(n, 1)};
    for (int i = 0; i < n; ++i)
      {
        const double li0 = L.value(i, 0.0);
        W.jump(i, 0)     = li0;
        for (int j = 0; j < n; ++j)
          {
            double m = 0.0, d = li0 * L.value(j, 0.0);
            for (size_t k = 0; k < q.x.size(); ++k)
              {
                m += q.w[k] * L.value(i, q.x[k]) * L.value(j, q.x[k]);
                d += q.w[k] * L.value(i, q.x[k]) * L.deriv(j, q.x[k]);
              }
            W.mass(i, j) = m;
            W.der(i, j)  = d;
          }
      }
    return W;
  }

  CGWeights get_cg_weights(int r)
  {
    const LagrangeBasis trial = get_time_basis(TimeStepType::CGP, r);
    LagrangeBasis       test;
    test.nodes.assign(trial.nodes.begin() + 1, trial.nodes.end());
    const Quad1D q = quad_gauss(r + 2);
    CGWeights    W{DMat(r, r + 1), DMat(r, r + 1)};
    for (int i = 0; i < r; ++i)
      for (int j = 0; j <= r; ++j)
        for (size_t k = 0; k < q.x.size(); ++k)
          {
            const double ti = q.w[k] * test.value(i, q.x[k]);
            W.mass(i, j) += ti * trial.value(j, q.x[k]);
            W.der(i, j) += ti * trial.deriv(j, q.x[k]);
          }
    return W;
  }

  TimeWeights get_fe_time_weights(TimeStepType type, int r, double tau, int nts)
  {
    // single-interval blocks: lhs mass/der, rhs parts (fe_time.h:300-311)
    DMat A1, B1, G1, Z1;
    if (type == TimeStepType::CGP)
      {
        CGWeights w = get_cg_weights(r);
        A1          = DMat(r, r);
        B1          = DMat(r, r);
        G1          = DMat(r, 1);
        Z1          = DMat(r, 1);
        for (int i = 0; i < r; ++i)
          {
            for (int j = 1; j <= r; ++j)
              {
                A1(i, j - 1) = tau * w.mass(i, j);
                B1(i, j - 1) = w.der(i, j);
              }
            G1(i, 0) = -tau * w.mass(i, 0);
            Z1(i, 0) = -w.der(i, 0);
          }
      }
    else
      {
        DGWeights w = get_dg_weights(r);
        A1          = w.mass;
        A1.scale(tau);
        B1 = w.der;
        G1 = DMat(r + 1, 1);   // no coupling through the K part
        Z1 = w.jump;           // jump term couples through M
      }
    const int   ni = A1.m, nt = ni * nts;
    TimeWeights W{DMat(nt, nt), DMat(nt, nt), DMat(nt, 1), DMat(nt, 1)};
    for (int it = 0; it < nts; ++it)
      {
        for (int i = 0; i < ni; ++i)
          for (int j = 0; j < ni; ++j)
            {
              W.Alpha(it * ni + i, it * ni + j) = A1(i, j);
              W.Beta(it * ni + i, it * ni + j)  = B1(i, j);
            }
        // coupling of step it+1 to the last dof of step it
        if (it + 1 < nts)
          for (int j = 0; j < ni; ++j)
            {
              W.Alpha((it + 1) * ni + j, it * ni + ni - 1) = -G1(j, 0);
              W.Beta((it + 1) * ni + j, it * ni + ni - 1)  = -Z1(j, 0);
            }
      }
    for (int i = 0; i < ni; ++i)
      {
        // Gamma couples the previous slab value through K, Zeta through M;
        // for DG the jump vector is returned as Gamma (fe_time.h:343-347)
        W.Gamma(i, 0) = (type == TimeStepType::CGP) ? G1(i, 0) : Z1(i, 0);
        W.Zeta(i, 0)  = (type == TimeStepType::CGP) ? Z1(i, 0) : G1(i, 0);
      }
    return W;
  }

  WaveWeights get_fe_time_weights_wave(TimeStepType type, const DMat &Alpha, const DMat &Beta,
                                       const DMat &Gamma, const DMat &Zeta, int nts)
  {
    const int  ni = Alpha.m, nt = ni * nts;
    const DMat Ai = Alpha.inverse();
    const DMat BAiB = Beta * (Ai * Beta);
    const DMat BAiG = Beta * (Ai * Gamma);
    const double aL = Alpha(ni - 1, ni - 1);
    const double gxai = Gamma(ni - 1, 0) / aL;
    DMat       GAiG = Gamma;
    GAiG.scale(gxai);
    DMat Brow(1, Beta.n);
    for (int j = 0; j < Beta.n; ++j)
      Brow(0, j) = Beta(Beta.m - 1, j);
    DMat GAiB = Gamma * Brow;
    GAiB.scale(1.0 / aL);

    WaveWeights W{DMat(nt, nt), DMat(nt, nt), DMat(nt, 1), DMat(nt, 1), DMat(nt, 1)};
    if (type == TimeStepType::CGP)
      {
        const DMat BAiZ = Beta * (Ai * Zeta);
        DMat       ZmG  = Zeta;
        for (size_t k = 0; k < ZmG.a.size(); ++k)
          ZmG.a[k] -= BAiG.a[k];
        DMat ZmB = ZmG * Brow;
        ZmB.scale(1.0 / aL);
        const double zxai = Zeta(ni - 1, 0) / aL;
        for (int it = 0; it < nts; ++it)
          for (int i = 0; i < ni; ++i)
            {
              const int row = it * ni + i;
              if (it == 0)
                {
                  W.rhs_uK(i, 0) = Gamma(i, 0);
                  W.rhs_uM(i, 0) = BAiZ(i, 0);
                  W.rhs_vM(i, 0) = ZmG(i, 0);
                }
              else
                {
                  W.rhs_uM(row, 0) = -zxai * std::pow(gxai, it - 1) * ZmG(i, 0);
                  W.rhs_vM(row, 0) = std::pow(gxai, it) * ZmG(i, 0);
                  W.lhs_uK(row, ni - 1 + (it - 1) * ni) = -Gamma(i, 0);
                  W.lhs_uM(row, ni - 1 + (it - 1) * ni) = -BAiZ(i, 0);
                }
              for (int j = 0; j < ni; ++j)
                {
                  W.lhs_uK(row, it * ni + j) = Alpha(i, j);
                  W.lhs_uM(row, it * ni + j) = BAiB(i, j);
                }
              for (int jt = 0; jt < it; ++jt)
                for (int j = 0; j < ni; ++j)
                  W.lhs_uM(row, jt * ni + j) +=
                    -std::pow(gxai, it - jt - 1) * ZmB(i, j) +
                    ((it > 1 && it - 1 > jt && j == ni - 1) ?
                       std::pow(gxai, it - jt - 2) * zxai * ZmG(i, 0) :
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
                  W.rhs_uM(i, 0) = BAiG(i, 0);
                  W.rhs_vM(i, 0) = Gamma(i, 0);
                }
              if (it == 1)
                W.rhs_uM(ni + i, 0) = -GAiG(i, 0);
              if (it + 1 < nts)
                for (int j = 0; j < ni; ++j)
                  W.lhs_uM(j + (it + 1) * ni, i + it * ni) =
                    -GAiB(j, i) - (i == ni - 1 ? BAiG(j, 0) : 0.0);
              if (it + 2 < nts && i == ni - 1)
                for (int j = 0; j < ni; ++j)
                  W.lhs_uM(j + (it + 2) * ni, i + it * ni) = GAiG(j, 0);
              for (int j = 0; j < ni; ++j)
                {
                  W.lhs_uK(i + it * ni, j + it * ni) = Alpha(i, j);
                  W.lhs_uM(i + it * ni, j + it * ni) = BAiB(i, j);
                }
            }
      }
    return W;
  }

  std::vector<TimeMGType> get_time_mg_sequence(int n_sp_lvl, int k_max, int k_min, int nts,
                                               int nts_min, bool time_before_space)
  {
    const int n_k = k_max - k_min;
    int       n_t = 0;
    for (int q = nts / nts_min; q > 1; q /= 2)
      ++n_t;
    std::vector<TimeMGType> seq(size_t(n_k + n_t + n_sp_lvl - 1), TimeMGType::none);
    // k levels sit lowest, then tau levels; with time_before_space both are
    // placed above the spatial levels
    const int k0 = time_before_space ? n_sp_lvl - 1 : 0;
    for (int i = 0; i < n_k; ++i)
      seq[k0 + i] = TimeMGType::k;
    for (int i = 0; i < n_t; ++i)
      seq[k0 + n_k + i] = TimeMGType::tau;
    return seq;
  }

  namespace
  {
    // interpolation of the parent 1D element onto child c's nodes
    DMat child_prolongation(const LagrangeBasis &L, int c)
    {
      const int n = L.size();
      DMat      E(n, n);
      for (int i = 0; i < n; ++i)
        {
          const double t = 0.5 * (L.nodes[i] + c);
          for (int j = 0; j < n; ++j)
            E(i, j) = L.value(j, t);
        }
      return E;
    }
    double integrate01(const std::function<double(double)> &f, double a, double b)
    {
      const Quad1D g = quad_gauss(16);
      double       s = 0;
      for (size_t q = 0; q < g.x.size(); ++q)
        s += g.w[q] * f(a + (b - a) * g.x[q]) * (b - a);
      return s;
    }
  } // namespace

  DMat get_time_prolongation_matrix(TimeStepType type, int r, int nts)
  {
    if (nts < 2 || (nts & (nts - 1)))
      throw std::runtime_error("time prolongation: n_timesteps_at_once must be a power of 2 >= 2");
    const LagrangeBasis L  = get_time_basis(type, r);
    const DMat          E0 = child_prolongation(L, 0), E1 = child_prolongation(L, 1);
    const int           ni = n_dofs_per_interval(type, r);
    // CGP drops the initial-value dof (row/col 0 of the lexicographic element)
    const int off = (type == TimeStepType::CGP) ? 1 : 0;
    DMat      P(ni * nts, ni * nts / 2);
    for (int it = 0; it < nts / 2; ++it)
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < ni; ++j)
          {
            P(2 * it * ni + i, it * ni + j)      = E0(i + off, j + off);
            P(2 * it * ni + ni + i, it * ni + j) = E1(i + off, j + off);
          }
    return P;
  }

  DMat get_time_restriction_matrix(TimeStepType type, int r, int nts)
  {
    if (nts < 2 || (nts & (nts - 1)))
      throw std::runtime_error("time restriction: n_timesteps_at_once must be a power of 2 >= 2");
    const LagrangeBasis L  = get_time_basis(type, r);
    const int           n  = L.size();
    const int           ni = n_dofs_per_interval(type, r);
    DMat                R0(n, n), R1(n, n);
    if (type == TimeStepType::DG)
      {
        // L2 projection of each child onto the parent space
        DMat Mp(n, n);
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            Mp(i, j) = integrate01([&](double t) { return L.value(i, t) * L.value(j, t); }, 0, 1);
        const DMat Mi = Mp.inverse();
        for (int c = 0; c < 2; ++c)
          {
            DMat X(n, n);
            for (int i = 0; i < n; ++i)
              for (int j = 0; j < n; ++j)
                X(i, j) = integrate01(
                  [&](double t) { return L.value(i, t) * L.value(j, 2.0 * t - c); }, 0.5 * c,
                  0.5 * c + 0.5);
            (c == 0 ? R0 : R1) = Mi * X;
          }
      }
    else
      {
        // interpolation: parent node values read from the child holding them
        for (int c = 0; c < 2; ++c)
          for (int i = 0; i < n; ++i)
            {
              const double x  = L.nodes[i];
              const bool   in = c == 0 ? x <= 0.5 + 1e-14 : x >= 0.5 - 1e-14;
              if (in)
                for (int j = 0; j < n; ++j)
                  (c == 0 ? R0 : R1)(i, j) = L.value(j, 2.0 * x - c);
            }
      }
    const int off = (type == TimeStepType::CGP) ? 1 : 0;
    DMat      R(ni * nts / 2, ni * nts);
    for (int it = 0; it < nts / 2; ++it)
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < ni; ++j)
          {
            R(it * ni + i, 2 * it * ni + j)      = R0(i + off, j + off);
            R(it * ni + i, 2 * it * ni + ni + j) = R1(i + off, j + off);
          }
    return R;
  }

  DMat get_time_projection_matrix(TimeStepType type, int r_src, int r_dst, int nts)
  {
    const LagrangeBasis Ls = get_time_basis(type, r_src), Ld = get_time_basis(type, r_dst);
    const int           ns = Ls.size(), nd = Ld.size();
    DMat                Md(nd, nd), X(nd, ns);
    for (int i = 0; i < nd; ++i)
      {
        for (int j = 0; j < nd; ++j)
          Md(i, j) = integrate01([&](double t) { return Ld.value(i, t) * Ld.value(j, t); }, 0, 1);
        for (int j = 0; j < ns; ++j)
          X(i, j) = integrate01([&](double t) { return Ld.value(i, t) * Ls.value(j, t); }, 0, 1);
      }
    const DMat P = Md.inverse() * X;
    if (type == TimeStepType::DG)
      {
        DMat Pn(nts * nd, nts * ns);
        for (int it = 0; it < nts; ++it)
          for (int i = 0; i < nd; ++i)
            for (int j = 0; j < ns; ++j)
              Pn(it * nd + i, it * ns + j) = P(i, j);
        return Pn;
      }
    // CGP: consecutive elements share the end/initial node; later steps
    // overwrite the shared row (block copy semantics of fe_time.h:612-614),
    // then the initial-value row and column are dropped (fe_time.h:616-621)
    const int di = r_dst, si = r_src;
    DMat      Full(nts * di + 1, nts * si + 1);
    for (int it = 0; it < nts; ++it)
      for (int i = 0; i < nd; ++i)
        for (int j = 0; j < ns; ++j)
          Full(it * di + i, it * si + j) = P(i, j);
    DMat Pn(nts * di, nts * si);
    for (int i = 1; i <= nts * di; ++i)
      for (int j = 1; j <= nts * si; ++j)
        Pn(i - 1, j - 1) = Full(i, j);
    return Pn;
  }
} // namespace stfem
