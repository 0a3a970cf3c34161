// Wavefront-local form of the fused matrix-free space-time operator in 3D
// (gfx950), used by k_operator.hip for the two-pass (cell buffer) operator:
//
//   dst_j = sum_i Alpha(j,i) K src_i + Beta(j,i) M src_i
//
// (SystemMatrix::vmult over MatrixFreeOperator K and M,
// tests/include/operators.h:94-122 and :487-548 of nlsmrg/dealii-stfem).
//
// A cell is handled by N^2 lanes of one group of wavefronts (a wavefront,
// four of them per workgroup never waiting for each other, or for N = 5, 6
// a whole workgroup; see WaveLayout).  For a 1D sweep along direction d
// every lane holds one whole line of N values along d, so the sweep is N^2 FMAs on registers with coefficients
// that are the same for all lanes (compile-time indices into the constant
// shape tables, i.e. scalar operands); between sweeps along different
// directions the lines are redistributed through the cell's LDS buffers.
// Where the group is one wavefront, only its lanes exchange data, and the
// LDS executes the instructions of a wavefront in issue order, so a
// wavefront-scope fence (which keeps the compiler from moving LDS accesses
// across it) replaces the workgroup barrier.
//
// Per input component i (at most four): values at the Gauss points by
// three sweeps (z, y, x), kept on the x lines.  Geometry once per cell (Q1
// Jacobian from the cell's vertices, affine along the x line): JxW and the
// six entries of the metric at every point, kept in registers.  Per output
// component j: V = sum_i Beta(j,i) q_i and U = sum_i Alpha(j,i) q_i at the
// Gauss points, the collocation gradient of U (x on the x lines, y and z on
// their own lines, gathered back to the x lines), H = G grad U, the
// transposed derivatives of H along each direction summed with JxW V on the
// x lines, and the transposed interpolation (x, y, z), which leaves the
// result on z lines for the cell-major stores.  Nothing is computed twice
// for different output components, and no per-output accumulator of all
// components is held, so the registers do not grow with nt.
#pragma once

#include "dev_operator.h"

namespace stfem
{
  namespace
  {
    // Keeps the compiler from moving LDS accesses of this wavefront across
    // this point (no instruction is emitted; the hardware already executes a
    // wavefront's LDS instructions in order).
    __device__ __forceinline__ void wave_sync()
    {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    // Cells are handled by groups of G wavefronts: G = 1 (LDS exchanges
    // ordered by wave_sync, four independent wavefronts per workgroup)
    // where N^2 lanes divide 64 well, G = 2 for N = 5 (five cells on 128
    // lanes instead of two on 64) and G = 4 for N = 6 (seven cells on 256
    // lanes instead of one on 64), the group being the workgroup and its
    // exchanges ordered by workgroup barriers.
    template <int N>
    struct WaveLayout
    {
      static constexpr int G     = N == 5 ? 2 : N == 6 ? 4 : 1; // wavefronts per group
      static constexpr int LG    = 64 * G;                     // lanes per group
      static constexpr int TPC   = N * N;                      // lanes per cell
      static constexpr int CPG   = LG / TPC;                   // cells per group
      static constexpr int BLOCK = G == 1 ? 256 : LG;          // workgroup size
      static constexpr int GPB   = BLOCK / LG;                 // groups per workgroup
      static constexpr int CPB   = CPG * GPB;                  // cells per workgroup
      static constexpr int NC    = N * N * N;                  // values per cell
      static constexpr int NB    = 4;                          // LDS line buffers per cell
      // LDS layout of a cell's buffer: point (x, y, z) at x + N y + SZ z,
      // buffers BS apart, cells PER apart (vertex coordinates after the
      // buffers).  The padding is chosen by a model of the bank conflicts
      // of the line accesses (32 banks of 4 bytes, 32 lanes of 4-byte and
      // 16 lanes of 8-byte accesses at a time): N = 4 halves the conflict
      // cycles of the unpadded layout, N = 3 and N = 5 take 30 % and 12 %
      // off.
      static constexpr int SZ    = N == 3 ? 9 : N == 4 ? 19 : N * N;
      static constexpr int BS    = N == 4 ? 76 : N * N * N;
      static constexpr int PER   = N == 3 ? 155 : N == 4 ? 336 : N == 5 ? 537 : NB * BS + 24;
      static_assert(BS >= (N - 1) * (1 + N + SZ) + 1 && PER >= NB * BS + 24, "LDS layout");
      static __device__ __forceinline__ void sync()
      {
        if constexpr (G == 1)
          wave_sync();
        else
          __syncthreads();
      }
    };

    // Lines of a cell in the three ownerships.  Lane (t0, t1) of the cell
    // holds the points (k, t0, t1) (OWN 0, lines along x), (t0, k, t1)
    // (OWN 1, along y) or (t0, t1, k) (OWN 2, along z), k = 0..N-1; a
    // buffer holds the cell's values at x + N y + SZ z (x fastest, z lines padded).
    template <int N, int OWN, int SZ = N * N>
    struct Line
    {
      static constexpr int stride = OWN == 0 ? 1 : OWN == 1 ? N : SZ;
      static __device__ __forceinline__ int base(int t0, int t1)
      {
        return OWN == 0 ? N * t0 + SZ * t1 : OWN == 1 ? t0 + SZ * t1 : t0 + N * t1;
      }
      template <typename T>
      static __device__ __forceinline__ void put(T *buf, int b, const T (&v)[N])
      {
#pragma unroll
        for (int k = 0; k < N; ++k)
          buf[b + k * stride] = v[k];
      }
      template <typename T>
      static __device__ __forceinline__ void get(const T *buf, int b, T (&v)[N])
      {
#pragma unroll
        for (int k = 0; k < N; ++k)
          v[k] = buf[b + k * stride];
      }
    };

    // y = A x for one of the four 1D kernels on a line (Gauss points q,
    // nodes c):  MAT 0 interp    y_q = sum_c S(q,c) x_c
    //            MAT 1 deriv     y_q = sum_m D(q,m) x_m
    //            MAT 2 interp_t  y_c = sum_q S(q,c) x_q
    //            MAT 3 deriv_t   y_m = sum_q D(q,m) x_q
    // N^2 FMAs, or in FP64 for N >= 5 the even-odd form (N = 5: 21
    // operations instead of 25, N = 6: 30 instead of 36; measured at 64^3
    // Q4: FP64 cell pass 6 % faster, FP32 - VALU-bound at four waves per
    // SIMD, where the N independent FMA chains of the plain form issue
    // back to back - 3 % slower, so FP32 keeps the plain form).
    template <int N, int MAT, typename T>
    __device__ __forceinline__ void kernel1d(const T (&x)[N], T (&y)[N])
    {
      using Tb = Tab<T>;
      if constexpr (N >= 5 && sizeof(T) == 8)
        {
          constexpr int  h     = N / 2;
          constexpr bool sym   = MAT % 2 == 0; // sigma = +1 for S, S^T
          T              e[h], o[h];
#pragma unroll
          for (int c = 0; c < h; ++c)
            {
              e[c] = x[c] + x[N - 1 - c];
              o[c] = x[c] - x[N - 1 - c];
            }
#pragma unroll
          for (int q = 0; q < h; ++q)
            {
              T P = Tb::EO(N, MAT, q * h) * e[0];
              T Q = Tb::EO(N, MAT, h * h + q * h) * o[0];
#pragma unroll
              for (int c = 1; c < h; ++c)
                {
                  P += Tb::EO(N, MAT, q * h + c) * e[c];
                  Q += Tb::EO(N, MAT, h * h + q * h + c) * o[c];
                }
              if constexpr (N % 2)
                P += Tb::EO(N, MAT, 2 * h * h + q) * x[h];
              y[q]         = P + Q;
              y[N - 1 - q] = sym ? P - Q : Q - P;
            }
          if constexpr (N % 2)
            {
              T m = sym ? Tb::EO(N, MAT, 2 * h * h + 2 * h) * x[h] : T(0);
#pragma unroll
              for (int c = 0; c < h; ++c)
                m += Tb::EO(N, MAT, 2 * h * h + h + c) * (sym ? e[c] : o[c]);
              y[h] = m;
            }
        }
      else
        {
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              auto A = [&](int c) {
                return MAT == 0 ? Tb::S(N, q, c) : MAT == 1 ? Tb::D(N, q, c)
                                                 : MAT == 2 ? Tb::S(N, c, q) : Tb::D(N, c, q);
              };
              T s = A(0) * x[0];
#pragma unroll
              for (int c = 1; c < N; ++c)
                s += A(c) * x[c];
              y[q] = s;
            }
        }
    }
    template <int N, typename T>
    __device__ __forceinline__ void interp(const T (&in)[N], T (&out)[N])
    {
      kernel1d<N, 0>(in, out);
    }
    template <int N, typename T>
    __device__ __forceinline__ void deriv(const T (&in)[N], T (&out)[N])
    {
      kernel1d<N, 1>(in, out);
    }
    template <int N, typename T>
    __device__ __forceinline__ void interp_t(const T (&in)[N], T (&out)[N])
    {
      kernel1d<N, 2>(in, out);
    }
    template <int N, typename T>
    __device__ __forceinline__ void deriv_t(const T (&in)[N], T (&out)[N])
    {
      kernel1d<N, 3>(in, out);
    }

    // Q1 Jacobian J[d][e] = dx_d / dxi_e of the cell with vertex
    // coordinates V[v * 3 + d] (vertex bits: x, y, z) at (xa, xb, xc).
    template <typename T>
    __device__ __forceinline__ void q1_jacobian(const T *V, T xa, T xb, T xc, T (&J)[3][3])
    {
      const T wx[2] = {T(1) - xa, xa}, wy[2] = {T(1) - xb, xb}, wz[2] = {T(1) - xc, xc};
#pragma unroll
      for (int d = 0; d < 3; ++d)
        {
          T s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
          for (int e = 0; e < 2; ++e)
#pragma unroll
            for (int f = 0; f < 2; ++f)
              {
                s0 += (V[(1 + 2 * e + 4 * f) * 3 + d] - V[(2 * e + 4 * f) * 3 + d]) * (wy[e] * wz[f]);
                s1 += (V[(e + 2 + 4 * f) * 3 + d] - V[(e + 4 * f) * 3 + d]) * (wx[e] * wz[f]);
                s2 += (V[(e + 2 * f + 4) * 3 + d] - V[(e + 2 * f) * 3 + d]) * (wx[e] * wy[f]);
              }
          J[d][0] = s0;
          J[d][1] = s1;
          J[d][2] = s2;
        }
    }

    // The operator on the cells of one workgroup (3D, two-pass form, NTI
    // input components): the results of output components [j0, jn) go to
    // the cell-major buffer `cells` (cell_slot layout, every local node).
    // Dirichlet nodes are read as zero.  lds: GPB * CPG * PER values.
    template <int N, int NTI, typename T>
    __device__ __forceinline__ void st_apply_wave(const StOpArgs<T> &A, const T *__restrict__ src,
                                                  T *__restrict__ cells, int j0, int jn, long blk,
                                                  T *__restrict__ lds)
    {
      using W = WaveLayout<N>;
      using LX = Line<N, 0, W::SZ>;
      using LY = Line<N, 1, W::SZ>;
      using LZ = Line<N, 2, W::SZ>;
      using Tb = Tab<T>;
      constexpr int TPC = W::TPC, CPG = W::CPG, NC = W::NC, p = N - 1;

      const int  grp    = int(threadIdx.x) / W::LG;
      const int  lane   = int(threadIdx.x) % W::LG;
      const bool active = lane < CPG * TPC; // lanes past the last cell slot
      const int  cs     = active ? lane / TPC : CPG - 1;
      const int  t      = lane % TPC;
      const int  t0 = t % N, t1 = t / N;
      const long cell  = (blk * W::GPB + grp) * CPG + cs;
      const bool valid = active && cell < A.g.n_cells;
      T *const   B0    = lds + (grp * CPG + cs) * W::PER;
      T *const   B1    = B0 + W::BS;
      T *const   B2    = B1 + W::BS;
      T *const   B3    = B2 + W::BS;
      T *const   Vl    = B3 + W::BS;
      const int  bx = LX::base(t0, t1), by = LY::base(t0, t1), bz = LZ::base(t0, t1);

      // --- cell position; the z line (t0, t1) of nodes and its Dirichlet mask
      const int nc0 = A.g.nc[0], nc1 = A.g.nc[1];
      int       cx = 0, cy = 0, cz = 0;
      if (valid)
        {
          const unsigned uc = unsigned(cell), u0 = unsigned(nc0);
          const unsigned r  = uc / u0;
          cx                = int(uc - r * u0);
          cy                = int(r % unsigned(nc1));
          cz                = int(r / unsigned(nc1));
        }
      const long nn0 = A.g.nn[0], nn1 = A.g.nn[1], nn2 = A.g.nn[2];
      const int  dir = A.g.dirichlet;
      unsigned   cmask = 0; // bit k: node k of the z line is a Dirichlet node
      {
        const long gx = long(cx) * p + t0, gy = long(cy) * p + t1;
        const bool bnd = (gx == 0 && (dir & 1)) || (gx == nn0 - 1 && (dir & 2)) ||
                         (gy == 0 && (dir & 4)) || (gy == nn1 - 1 && (dir & 8));
#pragma unroll
        for (int k = 0; k < N; ++k)
          {
            const long gz = long(cz) * p + k;
            if (bnd || (gz == 0 && (dir & 16)) || (gz == nn2 - 1 && (dir & 32)))
              cmask |= 1u << k;
          }
      }
      const T  *srcl = src + ((long(cx) * p + t0) + nn0 * ((long(cy) * p + t1) + nn1 * (long(cz) * p))) * NTI;
      const long ks  = nn0 * nn1 * NTI; // stride between the nodes of the z line

      // --- the cell's vertex coordinates (read by every lane of the cell
      //     in the geometry phase, after the syncs of the input phase)
      if (active)
        for (int e = t; e < 24; e += TPC)
          {
            const int v = e / 3, d = e - 3 * v;
            T         x = T((v >> d) & 1);
            if (valid)
              {
                const long nv0 = nc0 + 1, nv1 = nc1 + 1;
                const long vx = cx + (v & 1), vy = cy + ((v >> 1) & 1), vz = cz + ((v >> 2) & 1);
                x = T(A.vertices[(vx + nv0 * (vy + nv1 * vz)) * 3 + d]);
              }
            Vl[e] = x;
          }

      // --- values of every input component at the Gauss points, on x lines
      T q[NTI][N];
#pragma unroll
      for (int i = 0; i < NTI; ++i)
        {
          T a[N], b[N];
#pragma unroll
          for (int k = 0; k < N; ++k)
            a[k] = (valid && !((cmask >> k) & 1u)) ? srcl[k * ks + i] : T(0);
          interp<N>(a, b); // along z
          W::sync();
          if (active)
            LZ::put(B0, bz, b);
          W::sync();
          LY::get(B0, by, a);
          interp<N>(a, b); // along y
          if (active)
            LY::put(B1, by, b);
          W::sync();
          LX::get(B1, bx, a);
          interp<N>(a, q[i]); // along x
        }

      // --- geometry on the x line (Gauss points (k, t0, t1)): JxW and the
      //     metric G = kappa * w * adj(J) adj(J)^T / |det J| (symmetric, six
      //     entries), kept in registers for all output components; the
      //     Jacobian is affine in xi_x along the line, J = Jc + xi_x Jl
      //     (first column constant)
      T Gm[N][6], JxW[N];
      {
        const T xb = Tb::X(N, t0), xc = Tb::X(N, t1);
        T       Jc[3][3], Jl[3][3];
        q1_jacobian<T>(Vl, T(0), xb, xc, Jc);
        q1_jacobian<T>(Vl, T(1), xb, xc, Jl);
        const T wab = Tb::W(N, t0) * Tb::W(N, t1);
#pragma unroll
        for (int k = 0; k < N; ++k)
          {
            const T xa = Tb::X(N, k);
            T       J[3][3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
              {
                J[d][0] = Jc[d][0];
                J[d][1] = Jc[d][1] + xa * (Jl[d][1] - Jc[d][1]);
                J[d][2] = Jc[d][2] + xa * (Jl[d][2] - Jc[d][2]);
              }
            T adj[3][3];
            adj[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
            adj[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
            adj[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
            adj[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
            adj[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
            adj[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
            adj[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
            adj[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
            adj[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            const T det  = J[0][0] * adj[0][0] + J[0][1] * adj[1][0] + J[0][2] * adj[2][0];
            const T adet = det < T(0) ? -det : det;
            const T wq   = wab * Tb::W(N, k);
            T       kap  = A.coef_value;
            if (A.coef_mode == 1)
              kap = valid ? A.coef[cell] : T(0);
            else if (A.coef_mode == 2)
              kap = valid ? A.coef[cell * NC + k + N * t0 + N * N * t1] : T(0);
            JxW[k]     = adet * wq;
            const T gs = kap * wq / adet;
            int     m  = 0;
#pragma unroll
            for (int e = 0; e < 3; ++e)
#pragma unroll
              for (int f = e; f < 3; ++f)
                Gm[k][m++] = (adj[e][0] * adj[f][0] + adj[e][1] * adj[f][1] + adj[e][2] * adj[f][2]) * gs;
          }
      }

      // --- each output component j: V = sum_i Beta(j,i) q_i and
      //     U = sum_i Alpha(j,i) q_i at the Gauss points, the collocation
      //     gradient of U (x on the x lines, y and z on their lines, then
      //     gathered to the x lines), H = G grad U, r = JxW V + D_x^T H_x +
      //     D_y^T H_y + D_z^T H_z (summed on the x lines), then the
      //     transposed interpolation along x, y, z (result on z lines)
      const long L    = long(nc0) * N;
      const long cs_c = long(N) * L * A.nt_out; // cell_slot stride of the local z index
      T *const   dstl = cells + ((((long(cz) * nc1 + cy) * N) * N + t1) * L + long(cx) * N + t0) * A.nt_out;
      for (int j = j0; j < jn; ++j)
        {
          T V[N], U[N];
#pragma unroll
          for (int k = 0; k < N; ++k)
            V[k] = U[k] = T(0);
#pragma unroll
          for (int i = 0; i < NTI; ++i)
            {
              const T cb = A.beta[j * NTI + i];
              const T ca = A.alpha[j * NTI + i];
#pragma unroll
              for (int k = 0; k < N; ++k)
                {
                  V[k] += cb * q[i][k];
                  U[k] += ca * q[i][k];
                }
            }
          T gx[N], gy[N], gz[N], a[N];
          deriv<N>(U, gx);
          W::sync();
          if (active)
            LX::put(B0, bx, U);
          W::sync();
          LY::get(B0, by, a);
          deriv<N>(a, gy); // on y lines
          LZ::get(B0, bz, a);
          deriv<N>(a, gz); // on z lines
          if (active)
            {
              LY::put(B1, by, gy);
              LZ::put(B2, bz, gz);
            }
          W::sync();
          LX::get(B1, bx, gy);
          LX::get(B2, bx, gz);
          // metric (Hx, Hy, Hz in gx, gy, gz) and r = JxW V + D_x^T Hx
#pragma unroll
          for (int k = 0; k < N; ++k)
            {
              const T g0 = gx[k], g1 = gy[k], g2 = gz[k];
              gx[k] = Gm[k][0] * g0 + Gm[k][1] * g1 + Gm[k][2] * g2;
              gy[k] = Gm[k][1] * g0 + Gm[k][3] * g1 + Gm[k][4] * g2;
              gz[k] = Gm[k][2] * g0 + Gm[k][4] * g1 + Gm[k][5] * g2;
              V[k] *= JxW[k];
            }
          T r[N];
          deriv_t<N>(gx, r);
#pragma unroll
          for (int k = 0; k < N; ++k)
            r[k] += V[k];
          if (active)
            {
              LX::put(B3, bx, gy);
              LX::put(B0, bx, gz);
            }
          W::sync();
          LY::get(B3, by, a);
          deriv_t<N>(a, gy); // along y, on y lines
          LZ::get(B0, bz, a);
          deriv_t<N>(a, gz); // along z, on z lines
          if (active)
            {
              LY::put(B1, by, gy);
              LZ::put(B2, bz, gz);
            }
          W::sync();
          LX::get(B1, bx, gy);
          LX::get(B2, bx, gz);
#pragma unroll
          for (int k = 0; k < N; ++k)
            r[k] += gy[k] + gz[k];
          interp_t<N>(r, a); // along x
          if (active)
            LX::put(B3, bx, a);
          W::sync();
          LY::get(B3, by, r);
          interp_t<N>(r, a); // along y
          if (active)
            LY::put(B0, by, a);
          W::sync();
          LZ::get(B0, bz, r);
          interp_t<N>(r, a); // along z: nodes (t0, t1, k)
          if (valid)
            {
#pragma unroll
              for (int k = 0; k < N; ++k)
                dstl[k * cs_c + j] = a[k];
            }
        }
    }
  } // namespace
} // namespace stfem
