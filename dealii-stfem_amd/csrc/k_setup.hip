// Device setup of the factored cell-patch Vanka (gfx950).
//
// PreconditionVanka's constructor (tests/include/gmg.h:409-474 of
// nlsmrg/dealii-stfem) takes, for every cell c, the principal submatrices
// K[I_c,I_c] and M[I_c,I_c] of the globally assembled stiffness and mass
// matrices (constrained rows and columns reduced to their diagonal) and
// inverts B_c = Alpha (x) K_c + Beta (x) M_c.  The factored form used here
// needs, per cell, the generalized eigenpairs K_c S = M_c S Lambda with
// S^T M_c S = I and the small time blocks W_k = (lambda_k Alpha + Beta)^{-1}.
// The kernels below build the element matrices (same Gauss(p+1) rule, Q1
// mapping and coefficient as the operator), assemble the patch submatrices
// from the up to 3^dim cells around each patch, and pack the eigenvectors and
// the inverted time blocks after the batched eigensolver.
#include "kernels.h"

namespace stfem
{
  namespace
  {
    struct SetupTables
    {
      double S[MAX_N][MAX_N];  // phi_a(x_q), [q][a]
      double Sd[MAX_N][MAX_N]; // phi_a'(x_q)
      double qx[MAX_N], qw[MAX_N];
    };

    inline __device__ void cell_pos(const GridDesc &g, long c, int &cx, int &cy, int &cz)
    {
      cx     = int(c % g.nc[0]);
      long r = c / g.nc[0];
      if (g.dim == 3)
        {
          cy = int(r % g.nc[1]);
          cz = int(r / g.nc[1]);
        }
      else
        {
          cy = int(r);
          cz = 0;
        }
    }

    // Element matrices (row-major NL x NL, FP64) of cells c0 + blockIdx.x.
    template <int DIM, int N>
    __global__ void __launch_bounds__(256)
      k_element_matrices(const GridDesc g, const double *__restrict__ X, int coef_mode,
                         double coef_value, const double *__restrict__ coef, const SetupTables tb,
                         long c0, double *__restrict__ Me, double *__restrict__ Ke)
    {
      constexpr int NL = (DIM == 3) ? N * N * N : N * N, NQ = NL;
      __shared__ double jxw[NQ];
      __shared__ double G[NQ][6]; // kappa JxW J^{-1} J^{-T}, upper triangle
      // the 1D tables in LDS: the entry loop indexes them per lane, which
      // from the kernel-argument block would be a memory load per use
      __shared__ double Sl[N][N], Dl[N][N];
      for (int e = threadIdx.x; e < N * N; e += blockDim.x)
        {
          Sl[e / N][e % N] = tb.S[e / N][e % N];
          Dl[e / N][e % N] = tb.Sd[e / N][e % N];
        }
      const long cell = c0 + blockIdx.x;
      int        cx, cy, cz;
      cell_pos(g, cell, cx, cy, cz);
      const long nv0 = g.nc[0] + 1, nv1 = g.nc[1] + 1;
      for (int q = threadIdx.x; q < NQ; q += blockDim.x)
        {
          const int    qa = q % N, qb = (q / N) % N, qc = (DIM == 3) ? q / (N * N) : 0;
          const double xi[3] = {tb.qx[qa], tb.qx[qb], DIM == 3 ? tb.qx[qc] : 0.0};
          double       J[3][3] = {};
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
                    J[d][e] += gg * X[id * DIM + d];
                }
            }
          double Ji[3][3] = {}, det;
          if (DIM == 2)
            {
              det      = J[0][0] * J[1][1] - J[0][1] * J[1][0];
              Ji[0][0] = J[1][1] / det;
              Ji[0][1] = -J[0][1] / det;
              Ji[1][0] = -J[1][0] / det;
              Ji[1][1] = J[0][0] / det;
            }
          else
            {
              double a[3][3];
              a[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
              a[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
              a[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
              a[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
              a[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
              a[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
              a[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
              a[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
              a[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
              det     = J[0][0] * a[0][0] + J[0][1] * a[1][0] + J[0][2] * a[2][0];
              for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                  Ji[i][j] = a[i][j] / det;
            }
          const double w   = tb.qw[qa] * tb.qw[qb] * (DIM == 3 ? tb.qw[qc] : 1.0);
          const double JxW = fabs(det) * w;
          double       kap = coef_value;
          if (coef_mode == 1)
            kap = coef[cell];
          else if (coef_mode == 2)
            kap = coef[cell * NQ + q];
          jxw[q] = JxW;
          int k  = 0;
          for (int e = 0; e < 3; ++e)
            for (int f = e; f < 3; ++f, ++k)
              {
                double s = 0;
                for (int d = 0; d < DIM; ++d)
                  s += Ji[e][d] * Ji[f][d];
                G[q][k] = kap * JxW * s;
              }
        }
      __syncthreads();
      const size_t base = size_t(blockIdx.x) * NL * NL;
      for (int ent = threadIdx.x; ent < NL * NL; ent += blockDim.x)
        {
          const int i = ent / NL, j = ent - i * NL;
          const int ai = i % N, bi = (i / N) % N, ci = (DIM == 3) ? i / (N * N) : 0;
          const int aj = j % N, bj = (j / N) % N, cj = (DIM == 3) ? j / (N * N) : 0;
          double    m = 0, k = 0;
          for (int q = 0; q < NQ; ++q)
            {
              const int    qa = q % N, qb = (q / N) % N, qc = (DIM == 3) ? q / (N * N) : 0;
              const double sxi = Sl[qa][ai], syi = Sl[qb][bi];
              const double szi = (DIM == 3) ? Sl[qc][ci] : 1.0;
              const double sxj = Sl[qa][aj], syj = Sl[qb][bj];
              const double szj = (DIM == 3) ? Sl[qc][cj] : 1.0;
              const double gi[3] = {Dl[qa][ai] * syi * szi, sxi * Dl[qb][bi] * szi,
                                    DIM == 3 ? sxi * syi * Dl[qc][ci] : 0.0};
              const double gj[3] = {Dl[qa][aj] * syj * szj, sxj * Dl[qb][bj] * szj,
                                    DIM == 3 ? sxj * syj * Dl[qc][cj] : 0.0};
              m += sxi * syi * szi * sxj * syj * szj * jxw[q];
              const double *Gq = G[q];
              k += gi[0] * (Gq[0] * gj[0] + Gq[1] * gj[1] + Gq[2] * gj[2]) +
                   gi[1] * (Gq[1] * gj[0] + Gq[3] * gj[1] + Gq[4] * gj[2]) +
                   gi[2] * (Gq[2] * gj[0] + Gq[4] * gj[1] + Gq[5] * gj[2]);
            }
          Me[base + ent] = m;
          Ke[base + ent] = k;
        }
    }

    // Element matrices of 3D cells by sum factorization (row-major NL x NL,
    // FP64), cells c0 + blockIdx.x.  With phi_i = s(ai) s(bi) s(ci)
    // (s = 1D basis at the Gauss points, s' its derivative), the mass matrix
    // is sum_q phi_i phi_j JxW and the stiffness matrix sum_q sum_{d,e}
    // (d_d phi_i) G_de (d_e phi_j) with G = kappa JxW J^{-1} J^{-T}: ten
    // weighted fields f (the mass, and the nine (d, e) pairs) whose 1D
    // factors in direction x are s or s' on each side.  The quadrature sum
    // is contracted one direction at a time:
    //   T1[f][ai,aj][qb,qc] = sum_qa Fx(f,i)[qa][ai] Fx(f,j)[qa][aj] w_f[qa,qb,qc]
    //   T2[qc]              = sum_qb Fy(f,i)[qb][bi] Fy(f,j)[qb][bj] T1[f][ai,aj][qb,qc]
    //   A[ci][cj]          += sum_qc Fz(f,i)[qc][ci] Fz(f,j)[qc][cj] T2[qc]
    // (T2 per (ai, aj, bi, bj) in registers): about 10 (N^5 + N^6 + N^7)
    // multiply-adds per cell instead of N^9 of the entry-by-entry sums.
    // Dynamic LDS: 10 N^3 (weights) + 10 N^4 (T1) doubles.
    template <int N>
    __global__ void __launch_bounds__(256)
      k_element_matrices_sf(const GridDesc g, const double *__restrict__ X, int coef_mode,
                            double coef_value, const double *__restrict__ coef,
                            const SetupTables tb, long c0, double *__restrict__ Me,
                            double *__restrict__ Ke)
    {
      constexpr int NQ = N * N * N, NL = NQ, N2 = N * N, NF = 10;
      extern __shared__ double smem[];
      double *const wf = smem;           // [NF][NQ]
      double *const T1 = wf + NF * NQ;   // [NF][ai * N + aj][qb + N * qc]
      __shared__ double Sl[N][N], Dl[N][N];
      for (int e = threadIdx.x; e < N * N; e += blockDim.x)
        {
          Sl[e / N][e % N] = tb.S[e / N][e % N];
          Dl[e / N][e % N] = tb.Sd[e / N][e % N];
        }
      const long cell = c0 + blockIdx.x;
      int        cx, cy, cz;
      cell_pos(g, cell, cx, cy, cz);
      const long nv0 = g.nc[0] + 1, nv1 = g.nc[1] + 1;
      // weights at the Gauss points: JxW and kappa JxW J^{-1} J^{-T}
      for (int q = threadIdx.x; q < NQ; q += blockDim.x)
        {
          const int    qa = q % N, qb = (q / N) % N, qc = q / N2;
          const double xi[3] = {tb.qx[qa], tb.qx[qb], tb.qx[qc]};
          double       J[3][3] = {};
          for (int v = 0; v < 8; ++v)
            {
              const long id = (cx + (v & 1)) + nv0 * ((cy + ((v >> 1) & 1)) + nv1 * (cz + ((v >> 2) & 1)));
              for (int e = 0; e < 3; ++e)
                {
                  double gg = 1;
                  for (int f = 0; f < 3; ++f)
                    {
                      const int bit = (v >> f) & 1;
                      gg *= (f == e) ? (bit ? 1.0 : -1.0) : (bit ? xi[f] : 1 - xi[f]);
                    }
                  for (int d = 0; d < 3; ++d)
                    J[d][e] += gg * X[id * 3 + d];
                }
            }
          double a[3][3];
          a[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
          a[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
          a[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
          a[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
          a[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
          a[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
          a[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
          a[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
          a[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
          const double det = J[0][0] * a[0][0] + J[0][1] * a[1][0] + J[0][2] * a[2][0];
          double       Ji[3][3];
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              Ji[i][j] = a[i][j] / det;
          const double w   = tb.qw[qa] * tb.qw[qb] * tb.qw[qc];
          const double JxW = fabs(det) * w;
          double       kap = coef_value;
          if (coef_mode == 1)
            kap = coef[cell];
          else if (coef_mode == 2)
            kap = coef[cell * NQ + q];
          wf[q] = JxW;
          for (int d = 0; d < 3; ++d)
            for (int e = 0; e < 3; ++e)
              wf[(1 + 3 * d + e) * NQ + q] =
                kap * JxW * (Ji[d][0] * Ji[e][0] + Ji[d][1] * Ji[e][1] + Ji[d][2] * Ji[e][2]);
        }
      __syncthreads();
      // 1D factor of field f, side (0: i, 1: j), direction dir: the
      // derivative where the field's gradient component points along dir
      auto deriv_side = [](int f, int side, int dir) {
        if (f == 0)
          return false;
        const int d = (f - 1) / 3, e = (f - 1) % 3;
        return (side == 0 ? d : e) == dir;
      };
      // contraction over qa
      for (int idx = threadIdx.x; idx < NF * N2 * N2; idx += blockDim.x)
        {
          const int f = idx / (N2 * N2), r = idx - f * N2 * N2;
          const int ij = r / N2, bc = r - ij * N2;
          const int ai = ij / N, aj = ij - ai * N;
          const bool di = deriv_side(f, 0, 0), dj = deriv_side(f, 1, 0);
          double     acc = 0;
#pragma unroll
          for (int qa = 0; qa < N; ++qa)
            acc += (di ? Dl[qa][ai] : Sl[qa][ai]) * (dj ? Dl[qa][aj] : Sl[qa][aj]) *
                   wf[f * NQ + qa + N * bc];
          T1[idx] = acc;
        }
      __syncthreads();
      // contractions over qb (into registers) and qc, one (ai, aj, bi, bj,
      // ci) per thread and step: the row entries cj = 0..N-1
      const size_t base = size_t(blockIdx.x) * NL * NL;
      for (int tup = threadIdx.x; tup < N2 * N2 * N; tup += blockDim.x)
        {
          const int ij = tup / (N2 * N), rest = tup - ij * N2 * N;
          const int bij = rest / N, ci = rest - bij * N;
          const int ai = ij / N, aj = ij - ai * N, bi = bij / N, bj = bij - bi * N;
          double    m[N], k[N];
#pragma unroll
          for (int cj = 0; cj < N; ++cj)
            m[cj] = k[cj] = 0;
#pragma unroll 1
          for (int f = 0; f < NF; ++f)
            {
              const bool dyi = deriv_side(f, 0, 1), dyj = deriv_side(f, 1, 1);
              const bool dzi = deriv_side(f, 0, 2), dzj = deriv_side(f, 1, 2);
              const double(*Fyi)[N] = dyi ? Dl : Sl;
              const double(*Fyj)[N] = dyj ? Dl : Sl;
              const double(*Fzi)[N] = dzi ? Dl : Sl;
              const double(*Fzj)[N] = dzj ? Dl : Sl;
              const double *t1 = T1 + (f * N2 + ij) * N2;
              double        t2[N];
#pragma unroll
              for (int qc = 0; qc < N; ++qc)
                {
                  double acc = 0;
#pragma unroll
                  for (int qb = 0; qb < N; ++qb)
                    acc += Fyi[qb][bi] * Fyj[qb][bj] * t1[qb + N * qc];
                  t2[qc] = acc * Fzi[qc][ci];
                }
#pragma unroll
              for (int cj = 0; cj < N; ++cj)
                {
                  double acc = 0;
#pragma unroll
                  for (int qc = 0; qc < N; ++qc)
                    acc += Fzj[qc][cj] * t2[qc];
                  if (f == 0)
                    m[cj] = acc;
                  else
                    k[cj] += acc;
                }
            }
          const size_t row = base + size_t(ai + N * bi + N2 * ci) * NL + aj + N * bj;
#pragma unroll
          for (int cj = 0; cj < N; ++cj)
            {
              Me[row + N2 * cj] = m[cj];
              Ke[row + N2 * cj] = k[cj];
            }
        }
    }

    __device__ __forceinline__ bool constrained_node(const GridDesc &g, long gx, long gy, long gz)
    {
      const int d = g.dirichlet;
      bool c = (gx == 0 && (d & 1)) || (gx == g.nn[0] - 1 && (d & 2)) || (gy == 0 && (d & 4)) ||
               (gy == g.nn[1] - 1 && (d & 8));
      if (g.dim == 3)
        c = c || (gz == 0 && (d & 16)) || (gz == g.nn[2] - 1 && (d & 32));
      return c;
    }

    // Patch matrices of cells p0 + blockIdx.x, column-major NL x NL; the
    // element matrices of cells [e0, e0 + ne) are in Ke/Me (row-major).
    template <int DIM, int N>
    __global__ void __launch_bounds__(256)
      k_patch_matrices(const GridDesc g, long p0, long e0, const double *__restrict__ Me,
                       const double *__restrict__ Ke, double *__restrict__ Mp,
                       double *__restrict__ Kp)
    {
      constexpr int NL = (DIM == 3) ? N * N * N : N * N, p = N - 1;
      const long    cell = p0 + blockIdx.x;
      int           cx, cy, cz;
      cell_pos(g, cell, cx, cy, cz);
      const size_t base = size_t(blockIdx.x) * NL * NL;
      for (int ent = threadIdx.x; ent < NL * NL; ent += blockDim.x)
        {
          const int j = ent / NL, i = ent - j * NL; // column-major: i fastest
          const int li[3] = {i % N, (i / N) % N, (DIM == 3) ? i / (N * N) : 0};
          const int lj[3] = {j % N, (j / N) % N, (DIM == 3) ? j / (N * N) : 0};
          const int cc[3] = {cx, cy, cz};
          // neighbour offsets per direction that contain both nodes
          int lo[3], hi[3];
          for (int d = 0; d < 3; ++d)
            {
              lo[d] = 0;
              hi[d] = 0;
              if (d < DIM)
                {
                  if (li[d] == 0 && lj[d] == 0 && cc[d] > 0)
                    lo[d] = -1;
                  if (li[d] == p && lj[d] == p && cc[d] < g.nc[d] - 1)
                    hi[d] = 1;
                }
            }
          double m = 0, k = 0;
          for (int oz = lo[2]; oz <= hi[2]; ++oz)
            for (int oy = lo[1]; oy <= hi[1]; ++oy)
              for (int ox = lo[0]; ox <= hi[0]; ++ox)
                {
                  const long o = (cx + ox) + long(g.nc[0]) * ((cy + oy) + long(g.nc[1]) * (cz + oz));
                  const int  off[3] = {ox, oy, oz};
                  int        mi = 0, mj = 0, s = 1;
                  for (int d = 0; d < DIM; ++d)
                    {
                      mi += (li[d] - off[d] * p) * s;
                      mj += (lj[d] - off[d] * p) * s;
                      s *= N;
                    }
                  const size_t eb = size_t(o - e0) * NL * NL + size_t(mi) * NL + mj;
                  m += Me[eb];
                  k += Ke[eb];
                }
          if (i != j)
            {
              const long gxi = (long)cx * p + li[0], gyi = (long)cy * p + li[1], gzi = (long)cz * p + li[2];
              const long gxj = (long)cx * p + lj[0], gyj = (long)cy * p + lj[1], gzj = (long)cz * p + lj[2];
              if (constrained_node(g, gxi, gyi, gzi) || constrained_node(g, gxj, gyj, gzj))
                m = k = 0;
            }
          Mp[base + ent] = m;
          Kp[base + ent] = k;
        }
    }

    // S (row-major, eigenvector k in column k) and lambda_k in the level
    // precision for the cells of a chunk, with a check that the time block
    // lambda_k Alpha + Beta is invertible; Z: column-major eigenvectors, lam:
    // eigenvalues; one thread per (cell, k).
    template <typename T>
    __global__ void k_pack_factors(long ncell, int nl, int ld, int nt, const double *__restrict__ Z,
                                   const double *__restrict__ lam, const double *__restrict__ Alpha,
                                   const double *__restrict__ Beta, T *__restrict__ S,
                                   T *__restrict__ lam_out, int *__restrict__ singular)
    {
      const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (id >= ncell * nl)
        return;
      const long c = id / nl;
      const int  k = int(id - c * nl);
      const double *Zc = Z + size_t(c) * nl * nl;
      T            *Sc = S + size_t(c) * nl * ld;
      for (int i = 0; i < nl; ++i)
        Sc[size_t(i) * ld + k] = T(Zc[size_t(k) * nl + i]);
      if (k == 0) // row padding
        for (int i = 0; i < nl; ++i)
          for (int kk = nl; kk < ld; ++kk)
            Sc[size_t(i) * ld + kk] = T(0);
      const double l = lam[size_t(c) * nl + k];
      lam_out[size_t(c) * nl + k] = T(l);
      // the time block lambda_k Alpha + Beta must be invertible: elimination
      // with partial pivoting, a zero pivot flags the patch
      double A[8][8];
      for (int i = 0; i < nt; ++i)
        for (int j = 0; j < nt; ++j)
          A[i][j] = l * Alpha[i * nt + j] + Beta[i * nt + j];
      for (int col = 0; col < nt; ++col)
        {
          int    r    = col;
          double best = fabs(A[col][col]);
          for (int q = col + 1; q < nt; ++q)
            if (fabs(A[q][col]) > best)
              best = fabs(A[q][col]), r = q;
          if (best == 0.0)
            {
              atomicOr(singular, 1);
              return;
            }
          if (r != col)
            for (int j = 0; j < nt; ++j)
              {
                const double t = A[r][j];
                A[r][j]         = A[col][j];
                A[col][j]       = t;
              }
          for (int q = col + 1; q < nt; ++q)
            {
              const double f = A[q][col] / A[col][col];
              for (int j = col; j < nt; ++j)
                A[q][j] -= f * A[col][j];
            }
        }
    }

    template <int DIM, int N>
    void elements_dn(const GridDesc &g, const double *X, int coef_mode, double coef_value,
                     const double *coef, const SetupTables &tb, long c0, long nc, double *Me,
                     double *Ke, hipStream_t s)
    {
      if constexpr (DIM == 3)
        {
          constexpr size_t lds = size_t(10) * (N * N * N + N * N * N * N) * sizeof(double);
          if (lds > 64 * 1024)
            STFEM_HIP(hipFuncSetAttribute((const void *)k_element_matrices_sf<N>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
          k_element_matrices_sf<N><<<dim3((unsigned)nc), 256, lds, s>>>(g, X, coef_mode, coef_value,
                                                                        coef, tb, c0, Me, Ke);
        }
      else
        k_element_matrices<DIM, N><<<dim3((unsigned)nc), 256, 0, s>>>(g, X, coef_mode, coef_value,
                                                                      coef, tb, c0, Me, Ke);
    }
    template <int DIM, int N>
    void patches_dn(const GridDesc &g, long p0, long np, long e0, const double *Me,
                    const double *Ke, double *Mp, double *Kp, hipStream_t s)
    {
      k_patch_matrices<DIM, N><<<dim3((unsigned)np), 256, 0, s>>>(g, p0, e0, Me, Ke, Mp, Kp);
    }
  } // namespace

  void launch_element_matrices(const GridDesc &g, const double *X, int coef_mode,
                               double coef_value, const double *coef, const double *S,
                               const double *Sd, const double *qx, const double *qw, long c0,
                               long nc, double *Me, double *Ke, hipStream_t s)
  {
    const int N = g.p + 1;
    STFEM_REQUIRE(N >= 2 && N <= MAX_N, "element matrices: unsupported degree");
    if (nc == 0)
      return;
    SetupTables tb{};
    for (int q = 0; q < N; ++q)
      {
        tb.qx[q] = qx[q];
        tb.qw[q] = qw[q];
        for (int a = 0; a < N; ++a)
          {
            tb.S[q][a]  = S[q * N + a];
            tb.Sd[q][a] = Sd[q * N + a];
          }
      }
#define STFEM_EL(D, NN)                                                                           \
  case NN: elements_dn<D, NN>(g, X, coef_mode, coef_value, coef, tb, c0, nc, Me, Ke, s); break;
    if (g.dim == 2)
      switch (N)
        {
          STFEM_EL(2, 2) STFEM_EL(2, 3) STFEM_EL(2, 4) STFEM_EL(2, 5) STFEM_EL(2, 6)
        }
    else
      switch (N)
        {
          STFEM_EL(3, 2) STFEM_EL(3, 3) STFEM_EL(3, 4) STFEM_EL(3, 5) STFEM_EL(3, 6)
        }
#undef STFEM_EL
    STFEM_HIP(hipGetLastError());
  }

  void launch_patch_matrices(const GridDesc &g, long p0, long np, long e0, const double *Me,
                             const double *Ke, double *Mp, double *Kp, hipStream_t s)
  {
    const int N = g.p + 1;
    STFEM_REQUIRE(N >= 2 && N <= MAX_N, "patch matrices: unsupported degree");
    if (np == 0)
      return;
#define STFEM_PT(D, NN)                                                                           \
  case NN: patches_dn<D, NN>(g, p0, np, e0, Me, Ke, Mp, Kp, s); break;
    if (g.dim == 2)
      switch (N)
        {
          STFEM_PT(2, 2) STFEM_PT(2, 3) STFEM_PT(2, 4) STFEM_PT(2, 5) STFEM_PT(2, 6)
        }
    else
      switch (N)
        {
          STFEM_PT(3, 2) STFEM_PT(3, 3) STFEM_PT(3, 4) STFEM_PT(3, 5) STFEM_PT(3, 6)
        }
#undef STFEM_PT
    STFEM_HIP(hipGetLastError());
  }

  template <typename T>
  void launch_pack_factors(long ncell, int nl, int nt, const double *Z, const double *lam,
                           const double *Alpha, const double *Beta, T *S, T *lam_out,
                           int *singular, hipStream_t s)
  {
    STFEM_REQUIRE(nt >= 1 && nt <= 8, "pack factors: at most 8 time dofs");
    const long n = ncell * nl;
    if (n == 0)
      return;
    k_pack_factors<T><<<dim3((unsigned)((n + 127) / 128)), 128, 0, s>>>(
      ncell, nl, vanka_s_ld(nl, sizeof(T)), nt, Z, lam, Alpha, Beta, S, lam_out, singular);
    STFEM_HIP(hipGetLastError());
  }
  template void launch_pack_factors<double>(long, int, int, const double *, const double *,
                                            const double *, const double *, double *, double *,
                                            int *, hipStream_t);
  template void launch_pack_factors<float>(long, int, int, const double *, const double *,
                                           const double *, const double *, float *, float *,
                                           int *, hipStream_t);
} // namespace stfem
