// Device code of the fused matrix-free space-time operator (gfx950), used by
// the kernels of k_operator.hip.
//
//   dst_j = sum_i Alpha(j,i) K src_i + Beta(j,i) M src_i
//
// (SystemMatrix::vmult over MatrixFreeOperator K and M,
// tests/include/operators.h:94-122 and :487-548 of nlsmrg/dealii-stfem.)
// A cell is handled by N^(dim-1) threads, thread (a,b) owning the line of N
// values along the last coordinate direction (in registers); the other
// directions are swept through LDS.  All nt input components are evaluated
// once (values and reference gradients at the Gauss points); the time mixing
// with Alpha and Beta happens at the quadrature points, so each output
// component costs one integration.  Geometry (Q1 mapping, coefficient) is
// evaluated on the fly from the cell's 2^dim vertices.  Results are
// accumulated with hardware FP atomics; Dirichlet nodes are read as zero and
// never written.
//
// Every translation unit that includes this header owns its copy of the
// constant-memory shape tables and uploads it with upload_tables_here().
#pragma once

#include "dev_cells.h"

namespace stfem
{
  namespace
  {
    __constant__ double cSd[MAX_N + 1][MAX_N][MAX_N];
    __constant__ double cDd[MAX_N + 1][MAX_N][MAX_N];
    __constant__ double cWd[MAX_N + 1][MAX_N];
    __constant__ double cXd[MAX_N + 1][MAX_N];
    __constant__ float  cSf[MAX_N + 1][MAX_N][MAX_N];
    __constant__ float  cDf[MAX_N + 1][MAX_N][MAX_N];
    __constant__ float  cWf[MAX_N + 1][MAX_N];
    __constant__ float  cXf[MAX_N + 1][MAX_N];
    // Even-odd forms of the four 1D kernels (0: S, 1: D, 2: S^T, 3: D^T) in
    // FP64, matrices with A[q][c] = sigma A[N-1-q][N-1-c] (sigma = +1 for
    // S, -1 for D: both point sets are symmetric about the cell centre).
    // With h = N/2, e_c = x_c + x_{N-1-c}, o_c = x_c - x_{N-1-c} (c < h) and
    // m = x_h (odd N): y_q = P_q + Q_q and y_{N-1-q} = sigma (P_q - Q_q) for
    // q < h, P_q = sum_c E[q][c] e_c + M[q] m, Q_q = sum_c O[q][c] o_c,
    // E/O = (A[q][c] +/- A[q][N-1-c]) / 2, M[q] = A[q][h]; the middle row
    // y_h = sum_c R[c] (e_c or, for sigma = -1, o_c) + Rh m (sigma = +1).
    // Layout per matrix: E (h x h), O (h x h), M (h), R (h), Rh.
    constexpr int EO_LEN = 2 * (MAX_N / 2) * (MAX_N / 2) + 2 * (MAX_N / 2) + 1;
    __constant__ double cEOd[MAX_N + 1][4][EO_LEN];

    template <typename T>
    struct Tab;
    template <>
    struct Tab<double>
    {
      static __device__ __forceinline__ double S(int n, int q, int a) { return cSd[n][q][a]; }
      static __device__ __forceinline__ double D(int n, int q, int a) { return cDd[n][q][a]; }
      static __device__ __forceinline__ double W(int n, int q) { return cWd[n][q]; }
      static __device__ __forceinline__ double X(int n, int q) { return cXd[n][q]; }
      static __device__ __forceinline__ double EO(int n, int mat, int k) { return cEOd[n][mat][k]; }
    };
    template <>
    struct Tab<float>
    {
      static __device__ __forceinline__ float S(int n, int q, int a) { return cSf[n][q][a]; }
      static __device__ __forceinline__ float D(int n, int q, int a) { return cDf[n][q][a]; }
      static __device__ __forceinline__ float W(int n, int q) { return cWf[n][q]; }
      static __device__ __forceinline__ float X(int n, int q) { return cXf[n][q]; }
    };

    template <int DIM, int N>
    struct Layout
    {
      static constexpr int TPC = (DIM == 3) ? N * N : N;       // threads per cell
      static constexpr int CPB = 256 / TPC;                      // cells per block
      static constexpr int NC  = (DIM == 3) ? N * N * N : N * N; // values per cell
      static __device__ __forceinline__ int idx(int a, int b, int k)
      {
        return (DIM == 3) ? a + N * (b + N * k) : a + N * k;
      }
    };

    // One group of threads applies the operator to the CPB cells
    // [blk*CPB, blk*CPB + CPB) (cells past the end contribute nothing).
    // tid in [0, 256).  When the group has more than CPB*TPC threads (256
    // is not a multiple of TPC for N = 3, 5), the threads past CPB*TPC form
    // an idle slot CPB: they touch no global memory, and lds0/lds1 must then
    // hold (CPB+1)*NC values each so that their LDS traffic stays in that
    // slot (CPB*NC suffices for groups of exactly CPB*TPC threads).  Every
    // thread of the workgroup must call this with the same j0/jn: it
    // synchronises the workgroup a fixed number of times.
    // TO_CELLS: the cell results are stored into the cell-major buffer dst
    // (cell_slot layout, every local node including Dirichlet ones);
    // otherwise they are added to the node vector dst with atomics,
    // skipping Dirichlet nodes.
    // vlds: (1 << DIM) * DIM values per cell slot, the cell's vertex
    // coordinates (read by every thread of the cell in the geometry phase).
    template <int DIM, int N, int JB, typename T, bool TO_CELLS>
    __device__ __forceinline__ void st_apply_cells(const StOpArgs<T> &A, const T *__restrict__ src,
                                                   T *__restrict__ dst, int j0, int jn, long blk,
                                                   int tid, T *__restrict__ lds0,
                                                   T *__restrict__ lds1, T *__restrict__ vlds)
    {
      using Ly = Layout<DIM, N>;
      using Tb = Tab<T>;
      constexpr int TPC = Ly::TPC, CPB = Ly::CPB, NC = Ly::NC;

      const int  lc    = tid / TPC;
      const int  t     = tid - lc * TPC;
      const int  ia    = t % N;
      const int  ib    = (DIM == 3) ? t / N : 0;
      const long cell  = blk * CPB + lc;
      const bool valid = lc < CPB && cell < A.g.n_cells;
      T *const   L0    = lds0 + lc * NC;
      T *const   L1    = lds1 + lc * NC;
      constexpr int p  = N - 1;

      // --- cell position and the node line owned by this thread
      const int nc0 = A.g.nc[0], nc1 = A.g.nc[1];
      int       cx = 0, cy = 0, cz = 0;
      if (valid)
        {
          // 32-bit splits (levels have fewer than 2^31 cells)
          const unsigned uc = unsigned(cell), u0 = unsigned(nc0);
          const unsigned r  = uc / u0;
          cx                = int(uc - r * u0);
          if (DIM == 3)
            {
              cy = int(r % nc1);
              cz = int(r / nc1);
            }
          else
            cy = int(r);
        }
      const long nn0 = A.g.nn[0], nn1 = A.g.nn[1], nn2 = A.g.nn[2];
      const int  dir = A.g.dirichlet;
      long       base, kstride;
      unsigned   cmask = 0; // bit k: node k of the line is a Dirichlet node
      {
        const long gx = (long)cx * p + ia;
        bool       bnd = (gx == 0 && (dir & 1)) || (gx == nn0 - 1 && (dir & 2));
        if (DIM == 3)
          {
            const long gy = (long)cy * p + ib;
            bnd = bnd || (gy == 0 && (dir & 4)) || (gy == nn1 - 1 && (dir & 8));
            base    = gx + nn0 * (gy + nn1 * ((long)cz * p));
            kstride = nn0 * nn1;
            for (int k = 0; k < N; ++k)
              {
                const long gz = (long)cz * p + k;
                if (bnd || (gz == 0 && (dir & 16)) || (gz == nn2 - 1 && (dir & 32)))
                  cmask |= 1u << k;
              }
          }
        else
          {
            base    = gx + nn0 * ((long)cy * p);
            kstride = nn0;
            for (int k = 0; k < N; ++k)
              {
                const long gy = (long)cy * p + k;
                if (bnd || (gy == 0 && (dir & 4)) || (gy == nn1 - 1 && (dir & 8)))
                  cmask |= 1u << k;
              }
          }
      }
      const int nti = A.nt_in, nto = A.nt_out;
      // --- the cell's vertex coordinates into LDS (visible after the first
      //     barrier of the evaluation loop)
      constexpr int NVX = (1 << DIM) * DIM;
      T *const      Vl  = vlds + lc * NVX;
      if (lc < CPB)
        for (int e = t; e < NVX; e += TPC) // TPC may be smaller than NVX
          {
            const int v = e / DIM, d = e - v * DIM;
            T         x = T((v >> d) & 1);
            if (valid)
              {
                const long nv0 = nc0 + 1, nv1 = nc1 + 1;
                const long vx = cx + (v & 1), vy = cy + ((v >> 1) & 1);
                const long vz = (DIM == 3) ? cz + ((v >> 2) & 1) : 0;
                x = T(A.vertices[(vx + nv0 * (vy + nv1 * vz)) * DIM + d]);
              }
            Vl[e] = x;
          }

      // --- accumulators at the quadrature points of the owned line
      T accV[JB][N];
      T accG[JB][DIM][N];
#pragma unroll
      for (int jj = 0; jj < JB; ++jj)
#pragma unroll
        for (int k = 0; k < N; ++k)
          {
            accV[jj][k] = T(0);
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              accG[jj][d][k] = T(0);
          }

      for (int i = 0; i < nti; ++i)
        {
          T u[N];
#pragma unroll
          for (int k = 0; k < N; ++k)
            u[k] = (valid && !((cmask >> k) & 1u)) ? src[(base + k * kstride) * nti + i] : T(0);
          // interpolation along the owned line
          T w[N];
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              T s = T(0);
#pragma unroll
              for (int c = 0; c < N; ++c)
                s += Tb::S(N, q, c) * u[c];
              w[q] = s;
            }
          T qv[N], g0[N], g1[N], g2[N];
          __syncthreads();
          if constexpr (DIM == 3)
            {
#pragma unroll
              for (int k = 0; k < N; ++k)
                L0[Ly::idx(ia, ib, k)] = w[k];
              __syncthreads();
              T v[N];
#pragma unroll
              for (int k = 0; k < N; ++k)
                {
                  T s = T(0);
#pragma unroll
                  for (int b2 = 0; b2 < N; ++b2)
                    s += Tb::S(N, ib, b2) * L0[Ly::idx(ia, b2, k)];
                  v[k] = s;
                }
#pragma unroll
              for (int k = 0; k < N; ++k)
                L1[Ly::idx(ia, ib, k)] = v[k];
              __syncthreads();
#pragma unroll
              for (int k = 0; k < N; ++k)
                {
                  T s = T(0);
#pragma unroll
                  for (int a2 = 0; a2 < N; ++a2)
                    s += Tb::S(N, ia, a2) * L1[Ly::idx(a2, ib, k)];
                  qv[k] = s;
                }
#pragma unroll
              for (int k = 0; k < N; ++k)
                L0[Ly::idx(ia, ib, k)] = qv[k];
              __syncthreads();
#pragma unroll
              for (int k = 0; k < N; ++k)
                {
                  T sx = T(0), sy = T(0), sz = T(0);
#pragma unroll
                  for (int m = 0; m < N; ++m)
                    {
                      sx += Tb::D(N, ia, m) * L0[Ly::idx(m, ib, k)];
                      sy += Tb::D(N, ib, m) * L0[Ly::idx(ia, m, k)];
                      sz += Tb::D(N, k, m) * qv[m];
                    }
                  g0[k] = sx;
                  g1[k] = sy;
                  g2[k] = sz;
                }
            }
          else
            {
#pragma unroll
              for (int k = 0; k < N; ++k)
                L0[Ly::idx(ia, 0, k)] = w[k];
              __syncthreads();
#pragma unroll
              for (int k = 0; k < N; ++k)
                {
                  T s = T(0);
#pragma unroll
                  for (int a2 = 0; a2 < N; ++a2)
                    s += Tb::S(N, ia, a2) * L0[Ly::idx(a2, 0, k)];
                  qv[k] = s;
                }
#pragma unroll
              for (int k = 0; k < N; ++k)
                L1[Ly::idx(ia, 0, k)] = qv[k];
              __syncthreads();
#pragma unroll
              for (int k = 0; k < N; ++k)
                {
                  T sx = T(0), sy = T(0);
#pragma unroll
                  for (int m = 0; m < N; ++m)
                    {
                      sx += Tb::D(N, ia, m) * L1[Ly::idx(m, 0, k)];
                      sy += Tb::D(N, k, m) * qv[m];
                    }
                  g0[k] = sx;
                  g1[k] = sy;
                  g2[k] = T(0);
                }
            }
          // time mixing at the quadrature points
#pragma unroll
          for (int jj = 0; jj < JB; ++jj)
            {
              const int j = j0 + jj;
              if (j < jn)
                {
                  const T ca = A.alpha[j * nti + i];
                  const T cb = A.beta[j * nti + i];
#pragma unroll
                  for (int k = 0; k < N; ++k)
                    {
                      accV[jj][k] += cb * qv[k];
                      accG[jj][0][k] += ca * g0[k];
                      accG[jj][1][k] += ca * g1[k];
                      if constexpr (DIM == 3)
                        accG[jj][2][k] += ca * g2[k];
                    }
                }
            }
        }

      // --- geometry at the quadrature points: JxW and the metric
      //     G = kappa * w * adj(J) adj(J)^T / |det J|
      {
        // vertex v, coordinate d of the cell (LDS)
        auto Vx = [&](int v, int d) -> T { return Vl[v * DIM + d]; };
        const T xa = Tb::X(N, ia), xb = Tb::X(N, ib);
        // The Q1 Jacobian along the owned line (last reference coordinate
        // xi_l varying): the columns of the other directions are linear in
        // xi_l and the last column does not depend on it, so
        // J = Jc + xi_l Jl with Jc, Jl computed once per thread (FP64; in
        // FP32 the 15 extra live registers cost more than the arithmetic
        // they save, and J is formed from the vertices at every point).
        constexpr bool HOIST = sizeof(T) == 8;
        T              Jc[DIM][DIM], Jl[DIM][DIM];
        if constexpr (HOIST && DIM == 3)
          {
            const T wx[2] = {T(1) - xa, xa}, wy[2] = {T(1) - xb, xb};
#pragma unroll
            for (int d = 0; d < 3; ++d)
              {
                T z0 = T(0), z1 = T(0), y0 = T(0), y1 = T(0), s2 = T(0);
#pragma unroll
                for (int e = 0; e < 2; ++e)
                  {
                    z0 += (Vx(1 + 2 * e, d) - Vx(2 * e, d)) * wy[e];
                    z1 += (Vx(5 + 2 * e, d) - Vx(4 + 2 * e, d)) * wy[e];
                    y0 += (Vx(e + 2, d) - Vx(e, d)) * wx[e];
                    y1 += (Vx(e + 6, d) - Vx(e + 4, d)) * wx[e];
#pragma unroll
                    for (int f = 0; f < 2; ++f)
                      s2 += (Vx(e + 2 * f + 4, d) - Vx(e + 2 * f, d)) * (wx[e] * wy[f]);
                  }
                Jc[d][0] = z0;
                Jl[d][0] = z1 - z0;
                Jc[d][1] = y0;
                Jl[d][1] = y1 - y0;
                Jc[d][2] = s2;
                Jl[d][2] = T(0);
              }
          }
        else if constexpr (HOIST)
          {
#pragma unroll
            for (int d = 0; d < 2; ++d)
              {
                const T a0 = Vx(1, d) - Vx(0, d), a1 = Vx(3, d) - Vx(2, d);
                Jc[d][0] = a0;
                Jl[d][0] = a1 - a0;
                Jc[d][1] = (Vx(2, d) - Vx(0, d)) * (T(1) - xa) + (Vx(3, d) - Vx(1, d)) * xa;
                Jl[d][1] = T(0);
              }
          }
        const T wab = (DIM == 3) ? Tb::W(N, ia) * Tb::W(N, ib) : Tb::W(N, ia);
#pragma unroll
        for (int k = 0; k < N; ++k)
          {
            const T xc = Tb::X(N, k);
            T       J[DIM][DIM];
            if constexpr (HOIST)
              {
#pragma unroll
                for (int d = 0; d < DIM; ++d)
#pragma unroll
                  for (int e = 0; e < DIM; ++e)
                    J[d][e] = (e == DIM - 1) ? Jc[d][e] : Jc[d][e] + xc * Jl[d][e];
              }
            else if constexpr (DIM == 3)
              {
                const T wy[2] = {T(1) - xb, xb}, wz[2] = {T(1) - xc, xc}, wx[2] = {T(1) - xa, xa};
#pragma unroll
                for (int d = 0; d < 3; ++d)
                  {
                    T s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
                    for (int e = 0; e < 2; ++e)
#pragma unroll
                      for (int f = 0; f < 2; ++f)
                        {
                          s0 += (Vx(1 + 2 * e + 4 * f, d) - Vx(2 * e + 4 * f, d)) * wy[e] * wz[f];
                          s1 += (Vx(e + 2 + 4 * f, d) - Vx(e + 4 * f, d)) * wx[e] * wz[f];
                          s2 += (Vx(e + 2 * f + 4, d) - Vx(e + 2 * f, d)) * wx[e] * wy[f];
                        }
                    J[d][0] = s0;
                    J[d][1] = s1;
                    J[d][2] = s2;
                  }
              }
            else
              {
                const T wy[2] = {T(1) - xc, xc}, wx[2] = {T(1) - xa, xa};
#pragma unroll
                for (int d = 0; d < 2; ++d)
                  {
                    J[d][0] = (Vx(1, d) - Vx(0, d)) * wy[0] + (Vx(3, d) - Vx(2, d)) * wy[1];
                    J[d][1] = (Vx(2, d) - Vx(0, d)) * wx[0] + (Vx(3, d) - Vx(1, d)) * wx[1];
                  }
              }
            const T wq = wab * Tb::W(N, k);
            T adj[DIM][DIM];
            T det;
            if constexpr (DIM == 3)
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
                det = J[0][0] * adj[0][0] + J[0][1] * adj[1][0] + J[0][2] * adj[2][0];
              }
            else
              {
                adj[0][0] = J[1][1];
                adj[0][1] = -J[0][1];
                adj[1][0] = -J[1][0];
                adj[1][1] = J[0][0];
                det       = J[0][0] * J[1][1] - J[0][1] * J[1][0];
              }
            const T adet = det < T(0) ? -det : det;
            T       kap  = A.coef_value;
            if (A.coef_mode == 1)
              kap = valid ? A.coef[cell] : T(0);
            else if (A.coef_mode == 2)
              kap = valid ? A.coef[cell * NC + Ly::idx(ia, ib, k)] : T(0);
            const T JxW = adet * wq;
            const T gs  = kap * wq / adet;
            T       G[DIM][DIM]; // symmetric
#pragma unroll
            for (int e = 0; e < DIM; ++e)
#pragma unroll
              for (int f = e; f < DIM; ++f)
                {
                  T s = T(0);
#pragma unroll
                  for (int d = 0; d < DIM; ++d)
                    s += adj[e][d] * adj[f][d];
                  G[e][f] = s * gs;
                  G[f][e] = G[e][f];
                }
#pragma unroll
            for (int jj = 0; jj < JB; ++jj)
              {
                accV[jj][k] *= JxW;
                T gr[DIM];
#pragma unroll
                for (int e = 0; e < DIM; ++e)
                  {
                    T s = T(0);
#pragma unroll
                    for (int f = 0; f < DIM; ++f)
                      s += G[e][f] * accG[jj][f][k];
                    gr[e] = s;
                  }
#pragma unroll
                for (int e = 0; e < DIM; ++e)
                  accG[jj][e][k] = gr[e];
              }
          }
      }

      // --- integrate each output component and scatter
#pragma unroll
      for (int jj = 0; jj < JB; ++jj)
        {
          const int j = j0 + jj;
          if (j >= jn)
            break;
          T out[N];
          __syncthreads();
          if constexpr (DIM == 3)
            {
#pragma unroll
              for (int k = 0; k < N; ++k)
                {
                  L0[Ly::idx(ia, ib, k)] = accG[jj][0][k];
                  L1[Ly::idx(ia, ib, k)] = accG[jj][1][k];
                }
              __syncthreads();
              T r[N];
#pragma unroll
              for (int k = 0; k < N; ++k)
                {
                  T s = accV[jj][k];
#pragma unroll
                  for (int m = 0; m < N; ++m)
                    {
                      s += Tb::D(N, m, k) * accG[jj][2][m];
                      s += Tb::D(N, m, ia) * L0[Ly::idx(m, ib, k)];
                      s += Tb::D(N, m, ib) * L1[Ly::idx(ia, m, k)];
                    }
                  r[k] = s;
                }
              __syncthreads();
#pragma unroll
              for (int k = 0; k < N; ++k)
                L0[Ly::idx(ia, ib, k)] = r[k];
              __syncthreads();
              T s1[N];
#pragma unroll
              for (int k = 0; k < N; ++k)
                {
                  T s = T(0);
#pragma unroll
                  for (int m = 0; m < N; ++m)
                    s += Tb::S(N, m, ia) * L0[Ly::idx(m, ib, k)];
                  s1[k] = s;
                }
#pragma unroll
              for (int k = 0; k < N; ++k)
                L1[Ly::idx(ia, ib, k)] = s1[k];
              __syncthreads();
              T s2[N];
#pragma unroll
              for (int k = 0; k < N; ++k)
                {
                  T s = T(0);
#pragma unroll
                  for (int m = 0; m < N; ++m)
                    s += Tb::S(N, m, ib) * L1[Ly::idx(ia, m, k)];
                  s2[k] = s;
                }
#pragma unroll
              for (int c = 0; c < N; ++c)
                {
                  T s = T(0);
#pragma unroll
                  for (int k = 0; k < N; ++k)
                    s += Tb::S(N, k, c) * s2[k];
                  out[c] = s;
                }
            }
          else
            {
#pragma unroll
              for (int k = 0; k < N; ++k)
                L0[Ly::idx(ia, 0, k)] = accG[jj][0][k];
              __syncthreads();
              T r[N];
#pragma unroll
              for (int k = 0; k < N; ++k)
                {
                  T s = accV[jj][k];
#pragma unroll
                  for (int m = 0; m < N; ++m)
                    {
                      s += Tb::D(N, m, k) * accG[jj][1][m];
                      s += Tb::D(N, m, ia) * L0[Ly::idx(m, 0, k)];
                    }
                  r[k] = s;
                }
#pragma unroll
              for (int k = 0; k < N; ++k)
                L1[Ly::idx(ia, 0, k)] = r[k];
              __syncthreads();
              T s1[N];
#pragma unroll
              for (int k = 0; k < N; ++k)
                {
                  T s = T(0);
#pragma unroll
                  for (int m = 0; m < N; ++m)
                    s += Tb::S(N, m, ia) * L1[Ly::idx(m, 0, k)];
                  s1[k] = s;
                }
#pragma unroll
              for (int c = 0; c < N; ++c)
                {
                  T s = T(0);
#pragma unroll
                  for (int k = 0; k < N; ++k)
                    s += Tb::S(N, k, c) * s1[k];
                  out[c] = s;
                }
            }
          if (valid)
            {
              if constexpr (TO_CELLS)
                {
#pragma unroll
                  for (int c = 0; c < N; ++c)
                    {
                      const long slot = (DIM == 3) ? cell_slot<DIM, N>(A.g, nto, cx, cy, cz, ia, ib, c, j)
                                                   : cell_slot<DIM, N>(A.g, nto, cx, cy, cz, ia, c, 0, j);
                      dst[slot] = out[c];
                    }
                }
              else
                {
#pragma unroll
                  for (int c = 0; c < N; ++c)
                    if (!((cmask >> c) & 1u))
                      atomicAdd(&dst[(base + c * kstride) * nto + j], out[c]);
                }
            }
        }
    }

    // copy the 1D tables of FE_Q(N-1) into this translation unit's constant memory
    inline void upload_tables_here(const double *S, const double *D, const double *W,
                                   const double *X, int N)
    {
      double Sd[MAX_N][MAX_N] = {}, Dd[MAX_N][MAX_N] = {}, Wd[MAX_N] = {}, Xd[MAX_N] = {};
      float  Sf[MAX_N][MAX_N] = {}, Df[MAX_N][MAX_N] = {}, Wf[MAX_N] = {}, Xf[MAX_N] = {};
      for (int q = 0; q < N; ++q)
        {
          Wd[q] = W[q];
          Wf[q] = float(W[q]);
          Xd[q] = X[q];
          Xf[q] = float(X[q]);
          for (int a = 0; a < N; ++a)
            {
              Sd[q][a] = S[q * N + a];
              Dd[q][a] = D[q * N + a];
              Sf[q][a] = float(S[q * N + a]);
              Df[q][a] = float(D[q * N + a]);
            }
        }
      // even-odd coefficients of S, D, S^T, D^T
      double eod[4][EO_LEN] = {}, sd[MAX_N][MAX_N], dd[MAX_N][MAX_N];
      const int h = N / 2;
      for (int mat = 0; mat < 4; ++mat)
        {
          for (int q = 0; q < N; ++q)
            for (int c = 0; c < N; ++c)
              {
                const double *src = mat % 2 == 0 ? S : D;
                // A[q][c]: the matrix (S, D) or its transpose
                const double v = mat < 2 ? src[q * N + c] : src[c * N + q];
                (mat % 2 == 0 ? sd : dd)[q][c] = v;
              }
          const double(*A)[MAX_N] = mat % 2 == 0 ? sd : dd;
          double *e = eod[mat];
          for (int q = 0; q < h; ++q)
            for (int c = 0; c < h; ++c)
              {
                e[q * h + c]         = 0.5 * (A[q][c] + A[q][N - 1 - c]);
                e[h * h + q * h + c] = 0.5 * (A[q][c] - A[q][N - 1 - c]);
              }
          if (N % 2)
            {
              for (int q = 0; q < h; ++q)
                e[2 * h * h + q] = A[q][h];
              for (int c = 0; c < h; ++c)
                e[2 * h * h + h + c] = A[h][c];
              e[2 * h * h + 2 * h] = A[h][h];
            }
        }
      STFEM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cEOd), eod, sizeof(eod), size_t(N) * sizeof(eod),
                                  hipMemcpyHostToDevice));
      const size_t o2 = size_t(N) * sizeof(Sd), o1 = size_t(N) * sizeof(Wd);
      const size_t f2 = size_t(N) * sizeof(Sf), f1 = size_t(N) * sizeof(Wf);
      STFEM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cSd), Sd, sizeof(Sd), o2, hipMemcpyHostToDevice));
      STFEM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cDd), Dd, sizeof(Dd), o2, hipMemcpyHostToDevice));
      STFEM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cWd), Wd, sizeof(Wd), o1, hipMemcpyHostToDevice));
      STFEM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cXd), Xd, sizeof(Xd), o1, hipMemcpyHostToDevice));
      STFEM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cSf), Sf, sizeof(Sf), f2, hipMemcpyHostToDevice));
      STFEM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cDf), Df, sizeof(Df), f2, hipMemcpyHostToDevice));
      STFEM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cWf), Wf, sizeof(Wf), f1, hipMemcpyHostToDevice));
      STFEM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cXf), Xf, sizeof(Xf), f1, hipMemcpyHostToDevice));
    }
  } // namespace
} // namespace stfem
