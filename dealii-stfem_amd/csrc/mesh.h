// Mesh levels of the space-time slab solver on the host: FE_Q(p) 1D shape
// data, the structured level hierarchy of a subdivided, globally refined
// box with its geometry and coefficient, and the patch matrices of the
// Vanka setup.  Plain C++ (OpenMP), no HIP dependency.
#pragma once

#include <vector>

#include "grid.h"
#include "time_fe.h"

namespace stfem
{
  // ------------------------------------------------------------ geometry
  // FE_Q(p) on Gauss-Lobatto nodes with Gauss(p+1) quadrature.
  struct Shape1D
  {
    int                 p = 1, n = 2;
    std::vector<double> gll;            // nodes
    std::vector<double> qx, qw;         // Gauss points/weights
    std::vector<double> S, D, Sd;       // S[q][a], D[q][q'] (collocation), Sd[q][a] = phi_a'(x_q)
    LagrangeBasis       L;
    explicit Shape1D(int p_);
  };

  // One structured mesh level with its geometry and coefficient.
  struct MeshLevel
  {
    GridDesc            g{};
    std::vector<double> X;      // vertex coordinates
    int                 coef_mode = 0;
    double              coef_value = 1.0;
    std::vector<double> coef;   // per cell (mode 1) or per (cell, q) (mode 2)
  };

  // Builds the level hierarchy of a subdivided box refined `refine` times:
  // level s has subdiv * 2^s cells per direction and takes the finest
  // level's vertex positions (coarse vertices are a subset).  A per-cell
  // coefficient of the finest level is evaluated on coarser levels at the
  // quadrature points: coarse_coef 0 takes the value of the finest cell
  // containing the point (a coefficient function sampled at the points, as
  // MatrixFreeOperator::evaluate_coefficient does); 1, 2, 3 take the
  // geometric, harmonic or arithmetic mean over the finest cells that meet
  // the point's share of the coarse cell (per direction the interval between
  // the midpoints of neighbouring Gauss points), weighted by overlap.
  std::vector<MeshLevel> build_mesh_levels(int dim, const int subdiv[3], int refine, int p,
                                           const double *fine_vertices, const double *lo,
                                           const double *hi, const double *fine_kappa,
                                           int dirichlet_mask, int coarse_coef = 0);

  // Element mass / stiffness matrices of all cells (lexicographic local
  // numbering, n_c x n_c each), and their assembly into the principal
  // submatrices K[I_c, I_c], M[I_c, I_c] of the globally assembled matrices
  // with constrained rows/columns reduced to the diagonal.
  void cell_patch_matrices(const MeshLevel &m, const Shape1D &sh, std::vector<double> &Kp,
                           std::vector<double> &Mp);

} // namespace stfem
