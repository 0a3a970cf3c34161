// Per-cell coefficient gradient of the adjoint method (gfx950).
//
// The slab system A(kappa) x = R x_prev + f has the stiffness part
// Alpha (x) K(kappa) with K = sum_c kappa_c K_c (MatrixFreeOperator with a
// per-cell coefficient, tests/include/operators.h:445-457, :526-537 of
// nlsmrg/dealii-stfem).  With theta_c = log kappa_c,
//   dA/dtheta_c = kappa_c Alpha (x) K_c,
// so the misfit gradient contribution of one slab is
//   g_c -= kappa_c sum_j lambda_j^T K_c w_j,   w_j = sum_i Alpha(j,i) x_i
// (minus the rhs term when R carries K, for cG).  K_c is the unit-coefficient
// cell stiffness with the operator's Gauss(p+1) quadrature and Q1 geometry,
// so the gradient is that of the discrete operator to round-off.
// One workgroup per cell, one thread per quadrature point; the cell's node
// values of lambda and w sit in LDS.  Dirichlet nodes carry zero in both.
#include "kernels.h"

namespace stfem
{
  namespace
  {
    constexpr int GQ = MAX_N; // Gauss points per direction = p + 1

    struct GradSpace
    {
      double qw[GQ], qx[GQ];
      double S[GQ][MAX_N], D[GQ][MAX_N]; // phi_a and phi_a' at the Gauss points
    };

    template <int DIM, int N>
    __global__ void __launch_bounds__(256)
      k_cell_grad(const GridDesc g, const double *__restrict__ vertices, int nt,
                  const double *__restrict__ lam, const double *__restrict__ w,
                  const double *__restrict__ kappa, const GradSpace gs,
                  double *__restrict__ grad)
    {
      constexpr int NC = (DIM == 3) ? N * N * N : N * N, NQ = NC, p = N - 1;
      extern __shared__ double sm[];
      double *const lv = sm;            // [j][l]
      double *const wv = sm + nt * NC;  // [j][l]
      __shared__ double red[4];
      const long cell = blockIdx.x;
      const int  tid  = threadIdx.x;
      const int  cx = int(cell % g.nc[0]), cy = int((cell / g.nc[0]) % g.nc[1]);
      const int  cz = (DIM == 3) ? int(cell / ((long)g.nc[0] * g.nc[1])) : 0;
      const int  dmask = g.dirichlet;
      for (int e = tid; e < nt * NC; e += blockDim.x)
        {
          const int  j = e / NC, l = e - j * NC;
          const int  a = l % N, b = (l / N) % N, c = (DIM == 3) ? l / (N * N) : 0;
          const long gx = (long)cx * p + a, gy = (long)cy * p + b, gz = (long)cz * p + c;
          bool bnd = (gx == 0 && (dmask & 1)) || (gx == g.nn[0] - 1 && (dmask & 2)) ||
                     (gy == 0 && (dmask & 4)) || (gy == g.nn[1] - 1 && (dmask & 8));
          if (DIM == 3)
            bnd = bnd || (gz == 0 && (dmask & 16)) || (gz == g.nn[2] - 1 && (dmask & 32));
          const long node = gx + (long)g.nn[0] * (gy + (long)g.nn[1] * gz);
          lv[e] = bnd ? 0.0 : lam[node * nt + j];
          wv[e] = bnd ? 0.0 : w[node * nt + j];
        }
      __syncthreads();
      double acc = 0.0;
      if (tid < NQ)
        {
          const int    qa = tid % N, qb = (tid / N) % N, qc = (DIM == 3) ? tid / (N * N) : 0;
          const double xi[3] = {gs.qx[qa], gs.qx[qb], DIM == 3 ? gs.qx[qc] : 0.0};
          // Q1 Jacobian at the point
          double     J[3][3] = {};
          const long nv0 = g.nc[0] + 1, nv1 = g.nc[1] + 1;
          for (int v = 0; v < (1 << DIM); ++v)
            {
              const long id = (cx + (v & 1)) +
                              nv0 * ((cy + ((v >> 1) & 1)) + nv1 * (DIM == 3 ? cz + ((v >> 2) & 1) : 0));
              for (int e = 0; e < DIM; ++e)
                {
                  double gg = 1;
                  for (int f = 0; f < DIM; ++f)
                    {
                      const int bit = (v >> f) & 1;
                      gg *= (f == e) ? (bit ? 1.0 : -1.0) : (bit ? xi[f] : 1 - xi[f]);
                    }
                  for (int d = 0; d < DIM; ++d)
                    J[d][e] += gg * vertices[id * DIM + d];
                }
            }
          // G = w_q adj(J) adj(J)^T / |det J|: (grad_ref u)^T G (grad_ref v)
          // is w_q |det J| (J^-T grad_ref u) . (J^-T grad_ref v)
          double adj[3][3] = {}, det;
          if (DIM == 3)
            {
              adj[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
              adj[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
              adj[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
              adj[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
              adj[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
              adj[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
              adj[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
              adj[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
              adj[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
              det       = J[0][0] * adj[0][0] + J[0][1] * adj[1][0] + J[0][2] * adj[2][0];
            }
          else
            {
              adj[0][0] = J[1][1];
              adj[0][1] = -J[0][1];
              adj[1][0] = -J[1][0];
              adj[1][1] = J[0][0];
              det       = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            }
          const double wq = gs.qw[qa] * gs.qw[qb] * (DIM == 3 ? gs.qw[qc] : 1.0) / fabs(det);
          double       G[3][3];
          for (int e = 0; e < DIM; ++e)
            for (int f = 0; f < DIM; ++f)
              {
                double s = 0;
                for (int d = 0; d < DIM; ++d)
                  s += adj[e][d] * adj[f][d];
                G[e][f] = s * wq;
              }
          for (int j = 0; j < nt; ++j)
            {
              double gl[3] = {0, 0, 0}, gw[3] = {0, 0, 0};
              for (int l = 0; l < NC; ++l)
                {
                  const int    a = l % N, b = (l / N) % N, c = (DIM == 3) ? l / (N * N) : 0;
                  const double sx = gs.S[qa][a], sy = gs.S[qb][b];
                  const double sz = (DIM == 3) ? gs.S[qc][c] : 1.0;
                  const double r0 = gs.D[qa][a] * sy * sz, r1 = sx * gs.D[qb][b] * sz;
                  const double r2 = (DIM == 3) ? sx * sy * gs.D[qc][c] : 0.0;
                  const double vl = lv[j * NC + l], vw = wv[j * NC + l];
                  gl[0] += vl * r0;
                  gl[1] += vl * r1;
                  gl[2] += vl * r2;
                  gw[0] += vw * r0;
                  gw[1] += vw * r1;
                  gw[2] += vw * r2;
                }
              for (int e = 0; e < DIM; ++e)
                for (int f = 0; f < DIM; ++f)
                  acc += gl[e] * G[e][f] * gw[f];
            }
        }
      // workgroup sum (four wavefronts)
      for (int off = 32; off > 0; off >>= 1)
        acc += __shfl_down(acc, off, 64);
      if ((tid & 63) == 0)
        red[tid >> 6] = acc;
      __syncthreads();
      if (tid == 0)
        grad[cell] -= kappa[cell] * (red[0] + red[1] + red[2] + red[3]);
    }
  } // namespace

  void launch_cell_grad(const GridDesc &g, const double *vertices, int nt, const double *lam,
                        const double *w, const double *kappa, const double *qx, const double *qw,
                        const double *S, const double *D, double *grad, hipStream_t s)
  {
    const int N = g.p + 1;
    STFEM_REQUIRE(N >= 2 && N <= MAX_N, "cell gradient: unsupported degree");
    STFEM_REQUIRE((g.dim == 3 ? N * N * N : N * N) <= 256, "cell gradient: too many points");
    GradSpace gs{};
    for (int q = 0; q < N; ++q)
      {
        gs.qx[q] = qx[q];
        gs.qw[q] = qw[q];
        for (int a = 0; a < N; ++a)
          {
            gs.S[q][a] = S[q * N + a];
            gs.D[q][a] = D[q * N + a];
          }
      }
    const size_t lds  = 2 * size_t(nt) * (g.dim == 3 ? N * N * N : N * N) * sizeof(double);
    STFEM_REQUIRE(lds <= 64 * 1024, "cell gradient: too many time dofs");
    const dim3 grid{(unsigned)g.n_cells, 1, 1};
#define STFEM_GRAD_CASE(D_, N_)                                                                    \
  if (g.dim == D_ && N == N_)                                                                      \
    {                                                                                              \
      k_cell_grad<D_, N_><<<grid, 256, lds, s>>>(g, vertices, nt, lam, w, kappa, gs, grad);        \
      STFEM_HIP(hipGetLastError());                                                                \
      return;                                                                                      \
    }
    STFEM_GRAD_CASE(2, 2) STFEM_GRAD_CASE(2, 3) STFEM_GRAD_CASE(2, 4) STFEM_GRAD_CASE(2, 5)
    STFEM_GRAD_CASE(2, 6) STFEM_GRAD_CASE(3, 2) STFEM_GRAD_CASE(3, 3) STFEM_GRAD_CASE(3, 4)
    STFEM_GRAD_CASE(3, 5) STFEM_GRAD_CASE(3, 6)
#undef STFEM_GRAD_CASE
    throw Error(STATUS_ERR_UNSUPPORTED, "cell gradient: unsupported dim/degree");
  }
} // namespace stfem
