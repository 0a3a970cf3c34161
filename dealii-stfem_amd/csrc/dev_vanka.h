// Device code of the cell-patch Vanka smoother (gfx950), shared by the
// one-launch-per-level kernels and the coarse-level kernel.
//
// PreconditionVanka::vmult (tests/include/gmg.h:476-514 of
// nlsmrg/dealii-stfem), fused with the relaxation update of
// PreconditionRelaxation:
//
//   u += omega * sum_c (1/valence) .* scatter_c( B_c^{-1} gather_c(rhs - Au) )
//
// B_c = Alpha (x) K[I_c,I_c] + Beta (x) M[I_c,I_c], local index k + i*n_c
// (space fastest, time outer) as the reference builds it (gmg.h:460-468).
// Like the reference, the patch gathers and scatters every node of the cell,
// Dirichlet nodes included (their rows of B_c are decoupled).  The valence
// (number of cells sharing a node) of a structured-grid node is a product of
// per-direction multiplicities and is computed on the fly.
//
// Each function below handles one cell with one group of threads (tid in
// [0, nthr)); cells with valid == false touch no memory but pass the same
// workgroup barriers, so groups working on different cells of one workgroup
// stay in step.
#pragma once

#include "kernels.h"

namespace stfem
{
  namespace
  {
    struct PatchMap
    {
      int  dim, p, n; // n = p+1
      int  nc0, nc1;
      long nn0, nn1, nn2;

      __device__ void cell_xyz(long cell, int &cx, int &cy, int &cz) const
      {
        cx     = int(cell % nc0);
        long r = cell / nc0;
        if (dim == 3)
          {
            cy = int(r % nc1);
            cz = int(r / nc1);
          }
        else
          {
            cy = int(r);
            cz = 0;
          }
      }
      // global node of local node l (lexicographic within the cell) and its
      // inverse valence
      __device__ long node(int cx, int cy, int cz, int l, float &inv_val) const
      {
        const int  a  = l % n;
        const int  b  = (l / n) % n;
        const int  c  = (dim == 3) ? l / (n * n) : 0;
        const long gx = (long)cx * p + a, gy = (long)cy * p + b, gz = (long)cz * p + c;
        int m = ((gx % p) == 0 && gx > 0 && gx < nn0 - 1 ? 2 : 1) *
                ((gy % p) == 0 && gy > 0 && gy < nn1 - 1 ? 2 : 1);
        if (dim == 3)
          m *= ((gz % p) == 0 && gz > 0 && gz < nn2 - 1 ? 2 : 1);
        inv_val = 1.0f / float(m);
        return gx + nn0 * (gy + nn1 * gz);
      }
    };

    __host__ __device__ inline PatchMap make_map(const GridDesc &g)
    {
      PatchMap m;
      m.dim = g.dim;
      m.p   = g.p;
      m.n   = g.p + 1;
      m.nc0 = g.nc[0];
      m.nc1 = g.nc[1];
      m.nn0 = g.nn[0];
      m.nn1 = g.nn[1];
      m.nn2 = g.nn[2];
      return m;
    }

    inline int patch_size(const GridDesc &g)
    {
      return (g.dim == 3) ? (g.p + 1) * (g.p + 1) * (g.p + 1) : (g.p + 1) * (g.p + 1);
    }

    // xs[i*nloc + j] = (rhs - ax)[node_j][i]
    template <typename T>
    __device__ __forceinline__ void gather(const PatchMap &M, int cx, int cy, int cz, int nt,
                                           int nloc, const T *__restrict__ rhs,
                                           const T *__restrict__ ax, T *xs, int tid, int nthr)
    {
      const int nb = nt * nloc;
      for (int l = tid; l < nb; l += nthr)
        {
          const int  i = l / nloc, j = l - i * nloc;
          float      iv;
          const long g = M.node(cx, cy, cz, j, iv);
          T          v = rhs[g * nt + i];
          if (ax)
            v -= ax[g * nt + i];
          xs[l] = v;
        }
    }

    // Explicit inverse (column-major, a column read by consecutive lanes).
    // LDS: xs[nt*nloc].  One barrier.
    template <typename T>
    __device__ __forceinline__ void vanka_explicit_cell(const PatchMap &M, int nt, int nloc,
                                                        const T *__restrict__ binv,
                                                        const T *__restrict__ rhs,
                                                        const T *__restrict__ ax,
                                                        T *__restrict__ u, T omega, long cell,
                                                        bool valid, int tid, int nthr, T *xs)
    {
      const int nb = nt * nloc;
      int       cx = 0, cy = 0, cz = 0;
      if (valid)
        {
          M.cell_xyz(cell, cx, cy, cz);
          gather(M, cx, cy, cz, nt, nloc, rhs, ax, xs, tid, nthr);
        }
      __syncthreads();
      if (!valid)
        return;
      const T *B = binv + (size_t)cell * nb * nb;
      for (int row = tid; row < nb; row += nthr)
        {
          T acc = T(0);
          for (int col = 0; col < nb; ++col)
            acc += B[(size_t)col * nb + row] * xs[col];
          const int  i = row / nloc, j = row - i * nloc;
          float      iv;
          const long g = M.node(cx, cy, cz, j, iv);
          atomicAdd(&u[g * nt + i], omega * T(iv) * acc);
        }
    }

    // Factored apply for runtime nt and nloc (fallback shapes), one cell per
    // workgroup.  LDS: x, y (nt*nloc each) and, when STAGE, S (nloc*nloc).
    template <typename T, bool STAGE>
    __device__ __forceinline__ void vanka_factored_cell(const PatchMap &M, int nt, int nloc,
                                                        const T *__restrict__ S,
                                                        const T *__restrict__ W,
                                                        const T *__restrict__ rhs,
                                                        const T *__restrict__ ax,
                                                        T *__restrict__ u, T omega, long cell,
                                                        int tid, int nthr, T *smem)
    {
      T *const xs = smem;
      T *const ys = xs + nt * nloc;
      T *const ss = ys + nt * nloc;
      int      cx, cy, cz;
      M.cell_xyz(cell, cx, cy, cz);
      const T *Sg = S + (size_t)cell * nloc * nloc;
      if constexpr (STAGE)
        for (int l = tid; l < nloc * nloc; l += nthr)
          ss[l] = Sg[l];
      gather(M, cx, cy, cz, nt, nloc, rhs, ax, xs, tid, nthr);
      __syncthreads();
      const T *Sc = STAGE ? ss : Sg;
      const T *Wc = W + (size_t)cell * nloc * nt * nt;
      // y_{i,k} = sum_a S[a][k] x_{i,a}  then  z_{:,k} = W_k y_{:,k}
      for (int k = tid; k < nloc; k += nthr)
        {
          T yk[16];
          for (int i = 0; i < nt; ++i)
            yk[i] = T(0);
          for (int a = 0; a < nloc; ++a)
            {
              const T s = Sc[a * nloc + k];
              for (int i = 0; i < nt; ++i)
                yk[i] += s * xs[i * nloc + a];
            }
          const T *Wk = Wc + (size_t)k * nt * nt;
          for (int i = 0; i < nt; ++i)
            {
              T z = T(0);
              for (int j = 0; j < nt; ++j)
                z += Wk[i * nt + j] * yk[j];
              ys[i * nloc + k] = z;
            }
        }
      __syncthreads();
      // out_{i,a} = sum_k S[a][k] z_{i,k}
      for (int a = tid; a < nloc; a += nthr)
        {
          T o[16];
          for (int i = 0; i < nt; ++i)
            o[i] = T(0);
          const T *Sa = Sc + a * nloc;
          for (int k = 0; k < nloc; ++k)
            {
              const T s = Sa[k];
              for (int i = 0; i < nt; ++i)
                o[i] += s * ys[i * nloc + k];
            }
          float      iv;
          const long g = M.node(cx, cy, cz, a, iv);
          for (int i = 0; i < nt; ++i)
            atomicAdd(&u[g * nt + i], omega * T(iv) * o[i]);
        }
    }

    // Factored apply with compile-time patch size NL and time dofs NT: 256
    // threads per cell in P = 256/G groups of G = 64*ceil(NL/64) lanes.  Each
    // group handles a slice of the contraction index; partial sums meet in
    // LDS.  S is staged into LDS with a padded row stride (NL+1) so that both
    // the column reads of S^T x and the row reads of S z are conflict-free.
    // Four barriers.
    template <int NL>
    constexpr int v2_groups()
    {
      return 256 / (((NL + 63) / 64) * 64);
    }
    template <typename T, int NL, int NT>
    constexpr size_t v2_lds()
    {
      return sizeof(T) * (size_t(NL) * (NL + 1) + 2 * NT * NL + v2_groups<NL>() * NT * NL);
    }

    template <typename T, int NL, int NT>
    __device__ __forceinline__ void vanka_v2_cell(const PatchMap &M, const T *__restrict__ S,
                                                  const T *__restrict__ W,
                                                  const T *__restrict__ rhs,
                                                  const T *__restrict__ ax, T *__restrict__ u,
                                                  T omega, long cell, bool valid, int tid,
                                                  T *smem)
    {
      constexpr int G  = ((NL + 63) / 64) * 64;
      constexpr int P  = 256 / G;
      constexpr int LD = NL + 1;
      constexpr int CH = (NL + P - 1) / P; // contraction slice per group
      T *const ss  = smem;                 // NL * LD
      T *const xs  = ss + NL * LD;         // NT * NL
      T *const ys  = xs + NT * NL;         // NT * NL
      T *const red = ys + NT * NL;         // P * NT * NL

      int cx = 0, cy = 0, cz = 0;
      if (valid)
        {
          M.cell_xyz(cell, cx, cy, cz);
          const T *Sg = S + (size_t)cell * NL * NL;
          constexpr int VW = 16 / int(sizeof(T));
          if constexpr ((NL * NL) % VW == 0)
            {
              // 16-byte loads: every cell's block starts 16-byte aligned
              typedef T vec_t __attribute__((ext_vector_type(VW)));
              const vec_t *Sv = reinterpret_cast<const vec_t *>(Sg);
              for (int l4 = tid; l4 < NL * NL / VW; l4 += 256)
                {
                  const vec_t v = Sv[l4];
#pragma unroll
                  for (int e = 0; e < VW; ++e)
                    {
                      const int l = l4 * VW + e, a = l / NL, k = l - a * NL;
                      ss[a * LD + k] = v[e];
                    }
                }
            }
          else
            for (int l = tid; l < NL * NL; l += 256)
              {
                const int a = l / NL, k = l - a * NL;
                ss[a * LD + k] = Sg[l];
              }
          for (int l = tid; l < NT * NL; l += 256)
            {
              const int  i = l / NL, j = l - i * NL;
              float      iv;
              const long g = M.node(cx, cy, cz, j, iv);
              T          v = rhs[g * NT + i];
              if (ax)
                v -= ax[g * NT + i];
              xs[l] = v;
            }
        }
      __syncthreads();
      const int part = tid / G, lane = tid - part * G;
      const int a0 = part * CH, a1 = (a0 + CH < NL) ? a0 + CH : NL;
      // partial y_{i,k} over a in [a0, a1)
      if (valid && lane < NL)
        {
          T y[NT];
#pragma unroll
          for (int i = 0; i < NT; ++i)
            y[i] = T(0);
          for (int a = a0; a < a1; ++a)
            {
              const T s = ss[a * LD + lane];
#pragma unroll
              for (int i = 0; i < NT; ++i)
                y[i] += s * xs[i * NL + a];
            }
#pragma unroll
          for (int i = 0; i < NT; ++i)
            red[(part * NT + i) * NL + lane] = y[i];
        }
      __syncthreads();
      // z_{:,k} = W_k sum_parts y_{:,k}
      if (valid)
        for (int k = tid; k < NL; k += 256)
          {
            T y[NT];
#pragma unroll
            for (int i = 0; i < NT; ++i)
              {
                T s = T(0);
#pragma unroll
                for (int q = 0; q < P; ++q)
                  s += red[(q * NT + i) * NL + k];
                y[i] = s;
              }
            const T *Wk = W + ((size_t)cell * NL + k) * NT * NT;
#pragma unroll
            for (int i = 0; i < NT; ++i)
              {
                T z = T(0);
#pragma unroll
                for (int j = 0; j < NT; ++j)
                  z += Wk[i * NT + j] * y[j];
                ys[i * NL + k] = z;
              }
          }
      __syncthreads();
      // partial out_{i,a} over k in the group's slice
      if (valid && lane < NL)
        {
          T o[NT];
#pragma unroll
          for (int i = 0; i < NT; ++i)
            o[i] = T(0);
          for (int k = a0; k < a1; ++k)
            {
              const T s = ss[lane * LD + k];
#pragma unroll
              for (int i = 0; i < NT; ++i)
                o[i] += s * ys[i * NL + k];
            }
#pragma unroll
          for (int i = 0; i < NT; ++i)
            red[(part * NT + i) * NL + lane] = o[i];
        }
      __syncthreads();
      if (valid)
        for (int l = tid; l < NT * NL; l += 256)
          {
            const int i = l / NL, a = l - i * NL;
            T         s = T(0);
#pragma unroll
            for (int q = 0; q < P; ++q)
              s += red[(q * NT + i) * NL + a];
            float      iv;
            const long g = M.node(cx, cy, cz, a, iv);
            atomicAdd(&u[g * NT + i], omega * T(iv) * s);
          }
    }
  } // namespace
} // namespace stfem
