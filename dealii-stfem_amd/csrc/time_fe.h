// Time finite element algebra of the space-time slab solve (host, FP64).
//
// Mirrors the call surface of tests/include/fe_time.h of nlsmrg/dealii-stfem:
//   get_time_basis .............. fe_time.h:414-425
//   get_dg_weights / cg_weights . fe_time.h:473-555
//   split_lhs_rhs ............... fe_time.h:428-471
//   get_fe_time_weights ......... fe_time.h:292-381
//   get_fe_time_weights_wave .... fe_time.h:98-246, 383-412
//   get_time_mg_sequence ........ fe_time.h:43-69
//   prolongation / restriction / projection  fe_time.h:557-716
// Quadrature nodes come from eigenvalues of Jacobi matrices (Golub-Welsch).
#pragma once

#include <string>
#include <vector>

namespace stfem
{
  enum class TimeStepType : int
  {
    CGP = 1,
    DG  = 2
  };
  enum class TimeMGType : int
  {
    tau  = 1,
    k    = 2,
    none = 3
  };

  // Small dense row-major matrix.
  struct DMat
  {
    int                 m = 0, n = 0;
    std::vector<double> a;
    DMat() = default;
    DMat(int m_, int n_)
      : m(m_)
      , n(n_)
      , a(size_t(m_) * size_t(n_), 0.0)
    {}
    double &operator()(int i, int j) { return a[size_t(i) * n + j]; }
    double  operator()(int i, int j) const { return a[size_t(i) * n + j]; }
    bool    all_zero() const;
    DMat    transposed() const;
    DMat    operator*(const DMat &B) const;
    DMat    inverse() const;
    DMat   &scale(double s);
  };

  // Quadrature on [0,1].
  struct Quad1D
  {
    std::vector<double> x, w;
  };
  Quad1D quad_gauss(int n);
  Quad1D quad_gauss_lobatto(int n);
  Quad1D quad_gauss_radau_right(int n);

  // Lagrange polynomials on arbitrary nodes.
  struct LagrangeBasis
  {
    std::vector<double> nodes;
    int                 size() const { return int(nodes.size()); }
    double              value(int i, double t) const;
    double              deriv(int i, double t) const;
  };

  LagrangeBasis get_time_basis(TimeStepType type, int r);

  struct DGWeights
  {
    DMat mass, der, jump;
  };
  struct CGWeights
  {
    DMat mass, der;
  };
  DGWeights get_dg_weights(int r);
  CGWeights get_cg_weights(int r);

  // {Alpha, Beta, Gamma, Zeta} of fe_time.h:292-349
  struct TimeWeights
  {
    DMat Alpha, Beta, Gamma, Zeta;
  };
  TimeWeights get_fe_time_weights(TimeStepType type, int r, double tau, int n_timesteps_at_once);

  // {lhs_uK, lhs_uM, rhs_uK, rhs_uM, rhs_vM} of fe_time.h:98-246
  struct WaveWeights
  {
    DMat lhs_uK, lhs_uM, rhs_uK, rhs_uM, rhs_vM;
  };
  WaveWeights get_fe_time_weights_wave(TimeStepType type, const DMat &Alpha, const DMat &Beta,
                                       const DMat &Gamma, const DMat &Zeta,
                                       int n_timesteps_at_once = 1);

  std::vector<TimeMGType> get_time_mg_sequence(int n_sp_lvl, int k_max, int k_min,
                                               int n_timesteps_at_once,
                                               int n_timesteps_at_once_min, bool time_before_space);

  DMat get_time_prolongation_matrix(TimeStepType type, int r, int n_timesteps_at_once);
  DMat get_time_restriction_matrix(TimeStepType type, int r, int n_timesteps_at_once);
  DMat get_time_projection_matrix(TimeStepType type, int r_src, int r_dst,
                                  int n_timesteps_at_once);

  inline int n_dofs_per_interval(TimeStepType type, int r)
  {
    return type == TimeStepType::DG ? r + 1 : r;
  }
} // namespace stfem
