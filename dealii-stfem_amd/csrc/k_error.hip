// Space-time error of a slab against the manufactured heat solution
// u(t,x) = sin(2 pi f t) prod_d sin(2 pi f x_d) (gfx950).
//
// ErrorCalculator::evaluate_error (tests/include/exact_solution.h:218-299 of
// nlsmrg/dealii-stfem): at every Gauss point of every time step of the slab
// the numerical solution is the time-basis combination of the slab's blocks
// (and of the previous slab's last block for cG); the squared L2 and H1-
// seminorm errors are integrated with the cell Gauss rule and summed with the
// time weights, the L-infinity error is the maximum over the quadrature
// points.  One workgroup per cell, one thread per spatial quadrature point;
// the cell's node values of the current time point sit in LDS.
#include "kernels.h"

namespace stfem
{
  namespace
  {
    constexpr int EQ = 6; // largest number of Gauss points per direction

    struct ErrSpace
    {
      int    nq;
      double qx[EQ], qw[EQ];
      double S[EQ][MAX_N], D[EQ][MAX_N]; // phi_a and phi_a' of FE_Q(p) at the Gauss points
    };

    template <int DIM, int N>
    __global__ void __launch_bounds__(128)
      k_slab_error(const GridDesc g, const double *__restrict__ vertices, int nt,
                   const double *__restrict__ x, const double *__restrict__ prev, const ErrSpace es,
                   int npts, const double *__restrict__ tpts, double freq,
                   double *__restrict__ cell_err)
    {
      constexpr int NC = (DIM == 3) ? N * N * N : N * N, p = N - 1;
      __shared__ double u[NC];
      __shared__ double red[3][2];
      const long cell = blockIdx.x;
      const int  tid  = threadIdx.x;
      const int  cx = int(cell % g.nc[0]), cy = int((cell / g.nc[0]) % g.nc[1]);
      const int  cz = (DIM == 3) ? int(cell / ((long)g.nc[0] * g.nc[1])) : 0;
      const int  nq = es.nq, nqp = (DIM == 3) ? nq * nq * nq : nq * nq;
      const bool active = tid < nqp;
      const int  qa = tid % nq, qb = (tid / nq) % nq, qc = (DIM == 3) ? tid / (nq * nq) : 0;
      // geometry at this thread's point: x(xi), J, |det J|, J^{-1}
      double xq[3] = {0, 0, 0}, Ji[3][3] = {}, JxW = 0;
      if (active)
        {
          const double xi[3] = {es.qx[qa], es.qx[qb], DIM == 3 ? es.qx[qc] : 0.0};
          double       J[3][3] = {};
          const long   nv0 = g.nc[0] + 1, nv1 = g.nc[1] + 1;
          for (int v = 0; v < (1 << DIM); ++v)
            {
              const long id = (cx + (v & 1)) + nv0 * ((cy + ((v >> 1) & 1)) +
                                                      nv1 * (DIM == 3 ? cz + ((v >> 2) & 1) : 0));
              double N0 = 1;
              for (int e = 0; e < DIM; ++e)
                N0 *= ((v >> e) & 1) ? xi[e] : 1 - xi[e];
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
              for (int d = 0; d < DIM; ++d)
                xq[d] += N0 * vertices[id * DIM + d];
            }
          double det;
          if (DIM == 3)
            {
              const double a00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
              const double a01 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
              const double a02 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
              const double a10 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
              const double a11 = J[0][0] * J[2][2] - J[0][2] * J[2][0];
              const double a12 = J[0][2] * J[1][0] - J[0][0] * J[1][2];
              const double a20 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
              const double a21 = J[0][1] * J[2][0] - J[0][0] * J[2][1];
              const double a22 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
              det              = J[0][0] * a00 + J[0][1] * a10 + J[0][2] * a20;
              const double id  = 1.0 / det;
              Ji[0][0] = a00 * id, Ji[0][1] = a01 * id, Ji[0][2] = a02 * id;
              Ji[1][0] = a10 * id, Ji[1][1] = a11 * id, Ji[1][2] = a12 * id;
              Ji[2][0] = a20 * id, Ji[2][1] = a21 * id, Ji[2][2] = a22 * id;
            }
          else
            {
              det             = J[0][0] * J[1][1] - J[0][1] * J[1][0];
              const double id = 1.0 / det;
              Ji[0][0] = J[1][1] * id, Ji[0][1] = -J[0][1] * id;
              Ji[1][0] = -J[1][0] * id, Ji[1][1] = J[0][0] * id;
            }
          JxW = fabs(det) * es.qw[qa] * es.qw[qb] * (DIM == 3 ? es.qw[qc] : 1.0);
        }
      const double PI = 3.14159265358979323846;
      double       l2 = 0, h1 = 0, l8 = 0;
      const int    dmask = g.dirichlet;
      for (int tp = 0; tp < npts; ++tp)
        {
          // tpts row: time, weight, nt + 1 coefficients (blocks, then prev)
          const double *row = tpts + (size_t)tp * (nt + 3);
          for (int l = tid; l < NC; l += blockDim.x)
            {
              const int  a = l % N, b = (l / N) % N, c = (DIM == 3) ? l / (N * N) : 0;
              const long gx = (long)cx * p + a, gy = (long)cy * p + b, gz = (long)cz * p + c;
              bool bnd = (gx == 0 && (dmask & 1)) || (gx == g.nn[0] - 1 && (dmask & 2)) ||
                         (gy == 0 && (dmask & 4)) || (gy == g.nn[1] - 1 && (dmask & 8));
              if (DIM == 3)
                bnd = bnd || (gz == 0 && (dmask & 16)) || (gz == g.nn[2] - 1 && (dmask & 32));
              const long node = gx + (long)g.nn[0] * (gy + (long)g.nn[1] * gz);
              double     v    = 0.0;
              if (!bnd)
                {
                  for (int k = 0; k < nt; ++k)
                    v += row[2 + k] * x[node * nt + k];
                  v += row[2 + nt] * prev[node];
                }
              u[l] = v;
            }
          __syncthreads();
          if (active)
            {
              double uh = 0, gr[3] = {0, 0, 0};
              for (int l = 0; l < NC; ++l)
                {
                  const int    a = l % N, b = (l / N) % N, c = (DIM == 3) ? l / (N * N) : 0;
                  const double sx = es.S[qa][a], sy = es.S[qb][b];
                  const double sz = (DIM == 3) ? es.S[qc][c] : 1.0;
                  const double dx = es.D[qa][a], dy = es.D[qb][b];
                  const double dz = (DIM == 3) ? es.D[qc][c] : 0.0;
                  const double v  = u[l];
                  uh += v * sx * sy * sz;
                  gr[0] += v * dx * sy * sz;
                  gr[1] += v * sx * dy * sz;
                  gr[2] += v * sx * sy * dz;
                }
              // physical gradient: J^{-T} (reference gradient)
              double gp[3] = {0, 0, 0};
              for (int d = 0; d < DIM; ++d)
                for (int e = 0; e < DIM; ++e)
                  gp[d] += Ji[e][d] * gr[e];
              const double t  = row[0], w = row[1];
              const double st = sin(2 * PI * freq * t);
              double       sp = 1.0;
              for (int d = 0; d < DIM; ++d)
                sp *= sin(2 * PI * freq * xq[d]);
              const double e = uh - st * sp;
              l2 += w * JxW * e * e;
              l8 = fmax(l8, fabs(e));
              for (int d = 0; d < DIM; ++d)
                {
                  double ge = 2 * PI * freq * st;
                  for (int f = 0; f < DIM; ++f)
                    ge *= (f == d) ? cos(2 * PI * freq * xq[f]) : sin(2 * PI * freq * xq[f]);
                  h1 += w * JxW * (gp[d] - ge) * (gp[d] - ge);
                }
            }
          __syncthreads();
        }
      // workgroup reduction (two wavefronts)
      for (int off = 32; off > 0; off >>= 1)
        {
          l2 += __shfl_down(l2, off, 64);
          h1 += __shfl_down(h1, off, 64);
          l8 = fmax(l8, __shfl_down(l8, off, 64));
        }
      if ((tid & 63) == 0)
        {
          red[0][tid >> 6] = l2;
          red[1][tid >> 6] = h1;
          red[2][tid >> 6] = l8;
        }
      __syncthreads();
      if (tid == 0)
        {
          cell_err[cell * 3 + 0] = red[0][0] + red[0][1];
          cell_err[cell * 3 + 1] = fmax(red[2][0], red[2][1]);
          cell_err[cell * 3 + 2] = red[1][0] + red[1][1];
        }
    }
  } // namespace

  void launch_slab_error(const GridDesc &g, const double *vertices, int nt, const double *x,
                         const double *prev, int nq, const double *qx, const double *qw,
                         const double *S, const double *D, int npts, const double *tpts,
                         double freq, double *cell_err, hipStream_t s)
  {
    STFEM_REQUIRE(nq >= 1 && nq <= EQ, "slab error: unsupported number of Gauss points");
    const int N = g.p + 1;
    STFEM_REQUIRE(N >= 2 && N <= MAX_N, "slab error: unsupported degree");
    STFEM_REQUIRE((g.dim == 3 ? nq * nq * nq : nq * nq) <= 128,
                  "slab error: too many quadrature points per cell");
    ErrSpace es{};
    es.nq = nq;
    for (int q = 0; q < nq; ++q)
      {
        es.qx[q] = qx[q];
        es.qw[q] = qw[q];
        for (int a = 0; a < N; ++a)
          {
            es.S[q][a] = S[q * N + a];
            es.D[q][a] = D[q * N + a];
          }
      }
    const dim3 grid{(unsigned)g.n_cells, 1, 1};
    if (g.dim == 2)
      switch (N)
        {
          case 2: k_slab_error<2, 2><<<grid, 128, 0, s>>>(g, vertices, nt, x, prev, es, npts, tpts, freq, cell_err); break;
          case 3: k_slab_error<2, 3><<<grid, 128, 0, s>>>(g, vertices, nt, x, prev, es, npts, tpts, freq, cell_err); break;
          case 4: k_slab_error<2, 4><<<grid, 128, 0, s>>>(g, vertices, nt, x, prev, es, npts, tpts, freq, cell_err); break;
          case 5: k_slab_error<2, 5><<<grid, 128, 0, s>>>(g, vertices, nt, x, prev, es, npts, tpts, freq, cell_err); break;
          default: k_slab_error<2, 6><<<grid, 128, 0, s>>>(g, vertices, nt, x, prev, es, npts, tpts, freq, cell_err); break;
        }
    else
      switch (N)
        {
          case 2: k_slab_error<3, 2><<<grid, 128, 0, s>>>(g, vertices, nt, x, prev, es, npts, tpts, freq, cell_err); break;
          case 3: k_slab_error<3, 3><<<grid, 128, 0, s>>>(g, vertices, nt, x, prev, es, npts, tpts, freq, cell_err); break;
          case 4: k_slab_error<3, 4><<<grid, 128, 0, s>>>(g, vertices, nt, x, prev, es, npts, tpts, freq, cell_err); break;
          case 5: k_slab_error<3, 5><<<grid, 128, 0, s>>>(g, vertices, nt, x, prev, es, npts, tpts, freq, cell_err); break;
          default: k_slab_error<3, 6><<<grid, 128, 0, s>>>(g, vertices, nt, x, prev, es, npts, tpts, freq, cell_err); break;
        }
    STFEM_HIP(hipGetLastError());
  }
} // namespace stfem
