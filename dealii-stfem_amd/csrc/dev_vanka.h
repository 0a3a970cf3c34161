// Device code of the cell-patch Vanka smoother (gfx950), used by the
// kernels of k_vanka.hip.
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

#include "dev_cells.h"

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
        // 32-bit splits (levels have fewer than 2^31 cells)
        const unsigned uc = unsigned(cell), u0 = unsigned(nc0);
        const unsigned r  = uc / u0;
        cx                = int(uc - r * u0);
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

    // z = (lam Alpha + Beta)^{-1} y for one eigenmode of the factored
    // inverse (W_k applied without storing it): Gaussian elimination with
    // partial pivoting on the NT x NT time block, fully unrolled so that the
    // block stays in registers (row exchanges are selects).
    template <typename T, int NT>
    __device__ __forceinline__ void solve_time_block(T lam, const T *__restrict__ ta,
                                                     const T *__restrict__ tb, const T *y, T *z)
    {
      T M[NT][NT + 1];
#pragma unroll
      for (int i = 0; i < NT; ++i)
        {
#pragma unroll
          for (int j = 0; j < NT; ++j)
            M[i][j] = lam * ta[i * NT + j] + tb[i * NT + j];
          M[i][NT] = y[i];
        }
#pragma unroll
      for (int c = 0; c < NT; ++c)
        {
#pragma unroll
          for (int q = c + 1; q < NT; ++q)
            {
              const bool sw = fabs(M[q][c]) > fabs(M[c][c]);
#pragma unroll
              for (int j = c; j <= NT; ++j)
                {
                  const T a = M[c][j], b = M[q][j];
                  M[c][j] = sw ? b : a;
                  M[q][j] = sw ? a : b;
                }
            }
          const T inv = T(1) / M[c][c];
#pragma unroll
          for (int q = c + 1; q < NT; ++q)
            {
              const T f = M[q][c] * inv;
#pragma unroll
              for (int j = c + 1; j <= NT; ++j)
                M[q][j] -= f * M[c][j];
            }
        }
#pragma unroll
      for (int i = NT - 1; i >= 0; --i)
        {
          T s = M[i][NT];
#pragma unroll
          for (int j = i + 1; j < NT; ++j)
            s -= M[i][j] * z[j];
          z[i] = s / M[i][i];
        }
    }
    // the same for a run-time block size nt <= 16 (fallback shapes)
    template <typename T>
    __device__ inline void solve_time_block_rt(int nt, T lam, const T *__restrict__ ta,
                                               const T *__restrict__ tb, const T *y, T *z)
    {
      T M[16][17];
      for (int i = 0; i < nt; ++i)
        {
          for (int j = 0; j < nt; ++j)
            M[i][j] = lam * ta[i * nt + j] + tb[i * nt + j];
          M[i][nt] = y[i];
        }
      for (int c = 0; c < nt; ++c)
        {
          int r = c;
          for (int q = c + 1; q < nt; ++q)
            if (fabs(M[q][c]) > fabs(M[r][c]))
              r = q;
          if (r != c)
            for (int j = c; j <= nt; ++j)
              {
                const T t = M[c][j];
                M[c][j]   = M[r][j];
                M[r][j]   = t;
              }
          const T inv = T(1) / M[c][c];
          for (int q = c + 1; q < nt; ++q)
            {
              const T f = M[q][c] * inv;
              for (int j = c + 1; j <= nt; ++j)
                M[q][j] -= f * M[c][j];
            }
        }
      for (int i = nt - 1; i >= 0; --i)
        {
          T s = M[i][nt];
          for (int j = i + 1; j < nt; ++j)
            s -= M[i][j] * z[j];
          z[i] = s / M[i][i];
        }
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
    // S rows have the stride ld in global memory.
    template <typename T, bool STAGE>
    __device__ __forceinline__ void vanka_factored_cell(const PatchMap &M, int nt, int nloc,
                                                        int ld, const T *__restrict__ S,
                                                        const T *__restrict__ lam,
                                                        const T *__restrict__ ta,
                                                        const T *__restrict__ tb,
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
      const T *Sg = S + (size_t)cell * nloc * ld;
      if constexpr (STAGE)
        for (int l = tid; l < nloc * nloc; l += nthr)
          {
            const int a = l / nloc, k = l - a * nloc;
            ss[l]       = Sg[a * ld + k];
          }
      gather(M, cx, cy, cz, nt, nloc, rhs, ax, xs, tid, nthr);
      __syncthreads();
      const T  *Sc  = STAGE ? ss : Sg;
      const int sst = STAGE ? nloc : ld;
      const T *lc = lam + (size_t)cell * nloc;
      // y_{i,k} = sum_a S[a][k] x_{i,a}  then  z_{:,k} = W_k y_{:,k}
      for (int k = tid; k < nloc; k += nthr)
        {
          T yk[16];
          for (int i = 0; i < nt; ++i)
            yk[i] = T(0);
          for (int a = 0; a < nloc; ++a)
            {
              const T s = Sc[a * sst + k];
              for (int i = 0; i < nt; ++i)
                yk[i] += s * xs[i * nloc + a];
            }
          T zk[16];
          solve_time_block_rt<T>(nt, lc[k], ta, tb, yk, zk);
          for (int i = 0; i < nt; ++i)
            ys[i * nloc + k] = zk[i];
        }
      __syncthreads();
      // out_{i,a} = sum_k S[a][k] z_{i,k}
      for (int a = tid; a < nloc; a += nthr)
        {
          T o[16];
          for (int i = 0; i < nt; ++i)
            o[i] = T(0);
          const T *Sa = Sc + a * sst;
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

    // Factored apply with compile-time patch size NL and time dofs NT: one
    // cell per 256-thread workgroup, in P = 256/G groups of G = 64*ceil(NL/64)
    // lanes; each group handles a slice of the contraction index and the
    // partial sums meet in LDS.  S (row a = contraction index of S^T x) is
    // stored with rows padded to 16 bytes and copied to LDS with 16-byte
    // loads, all of a thread's loads issued before its first LDS store; the
    // LDS row stride is odd (NL, or NL+1 for even NL) so that the strided
    // row reads of S z are conflict-free.  The eigenvalue lambda_k is loaded
    // with S; W_k = (lambda_k Alpha + Beta)^{-1} is applied by solving the
    // time block.  The
    // local vectors are node-major ([a][i]) so a node's NT values are one
    // broadcast read.  Four barriers.
    // TWO_PASS: the corrections omega/valence * B_c^{-1} r_c go to the
    // cell-major buffer `cells` (cell_slot layout, NT components) for the
    // node gather; otherwise they are added to u with atomics.
    template <int NL>
    struct V2Shape
    {
      // patch sizes N^dim with N <= MAX_N determine dim and N
      static constexpr int dim = (NL == 8 || NL == 27 || NL == 64 || NL == 125 || NL == 216) ? 3 : 2;
      static constexpr int n   = dim == 3 ? (NL == 8 ? 2 : NL == 27 ? 3 : NL == 64 ? 4 : NL == 125 ? 5 : 6)
                                          : (NL == 4 ? 2 : NL == 9 ? 3 : NL == 16 ? 4 : NL == 25 ? 5 : 6);
      static_assert((dim == 3 ? n * n * n : n * n) == NL, "unsupported patch size");
    };
    // global node and inverse valence of local node a, patch shape known at
    // compile time (the divisions by n become multiplications)
    template <int NL>
    __device__ __forceinline__ long node_ct(const PatchMap &M, int cx, int cy, int cz, int a,
                                            float &inv_val)
    {
      using Sh         = V2Shape<NL>;
      constexpr int n  = Sh::n, p = n - 1;
      const int     la = a % n, lb = (a / n) % n, lc = (Sh::dim == 3) ? a / (n * n) : 0;
      const long    gx = (long)cx * p + la, gy = (long)cy * p + lb, gz = (long)cz * p + lc;
      int m = ((la == 0 && gx > 0) || (la == p && gx < M.nn0 - 1) ? 2 : 1) *
              ((lb == 0 && gy > 0) || (lb == p && gy < M.nn1 - 1) ? 2 : 1);
      if (Sh::dim == 3)
        m *= ((lc == 0 && gz > 0) || (lc == p && gz < M.nn2 - 1) ? 2 : 1);
      inv_val = 1.0f / float(m);
      return gx + M.nn0 * (gy + M.nn1 * gz);
    }

    template <int NL>
    constexpr int v2_groups()
    {
      return 256 / (((NL + 63) / 64) * 64);
    }
    template <int NL>
    constexpr int v2_ld()
    {
      return (NL & 1) ? NL : NL + 1;
    }
    // elements of the LDS region of S, rounded to 16 bytes
    template <typename T, int NL>
    constexpr int v2_sreg()
    {
      constexpr int VW = 16 / int(sizeof(T));
      return (NL * v2_ld<NL>() + VW - 1) / VW * VW;
    }
    template <typename T, int NL, int NT>
    constexpr size_t v2_lds()
    {
      return sizeof(T) * (size_t(v2_sreg<T, NL>()) + 2 * NT * NL + v2_groups<NL>() * NT * NL);
    }

    template <typename T, int NL, int NT, bool TWO_PASS>
    __device__ __forceinline__ void vanka_v2_cell(const PatchMap &M, const T *__restrict__ S,
                                                  const T *__restrict__ lam,
                                                  const T *__restrict__ ta,
                                                  const T *__restrict__ tb,
                                                  const T *__restrict__ rhs,
                                                  const T *__restrict__ ax, T *__restrict__ u,
                                                  T *__restrict__ cells, T omega, long cell,
                                                  int tid, T *smem)
    {
      constexpr int G  = ((NL + 63) / 64) * 64;
      constexpr int P  = 256 / G;
      constexpr int LD = v2_ld<NL>();
      constexpr int CH = (NL + P - 1) / P; // contraction slice per group
      constexpr int VW = 16 / int(sizeof(T));
      typedef T vec_t __attribute__((ext_vector_type(VW)));
      T *const xs  = smem + v2_sreg<T, NL>(); // NL x NT
      T *const ys  = xs + NT * NL;            // NL x NT
      T *const red = ys + NT * NL;            // P x NT x NL

      int cx, cy, cz;
      M.cell_xyz(cell, cx, cy, cz);
      // eigenvalue of this thread's mode k
      const T lk = (tid < NL) ? lam[(size_t)cell * NL + tid] : T(0);
      // S into LDS: 16-byte loads of the row-padded block (row stride LDG),
      // all of a thread's loads issued before its first LDS store
      {
        constexpr int LDG = (NL + VW - 1) / VW * VW;
        constexpr int NV  = NL * LDG / VW;
        constexpr int PER = (NV + 255) / 256;
        constexpr int RB  = PER < 16 ? PER : 16;
        const vec_t  *src = reinterpret_cast<const vec_t *>(S + (size_t)cell * NL * LDG);
        for (int r0 = 0; r0 < PER; r0 += RB)
          {
            vec_t v[RB];
#pragma unroll
            for (int m = 0; m < RB; ++m)
              {
                const int i = (r0 + m) * 256 + tid;
                if (r0 + m < PER && i < NV)
                  v[m] = src[i];
              }
#pragma unroll
            for (int m = 0; m < RB; ++m)
              {
                const int i = (r0 + m) * 256 + tid;
                if (r0 + m < PER && i < NV)
#pragma unroll
                  for (int e = 0; e < VW; ++e)
                    {
                      const int l = i * VW + e, a = l / LDG, k = l - a * LDG;
                      if (k < NL)
                        smem[a * LD + k] = v[m][e];
                    }
              }
          }
      }
      const T *ss = smem;
      // local residual, node-major
      for (int l = tid; l < NT * NL; l += 256)
        {
          const int  a = l / NT, i = l - a * NT;
          float      iv;
          const long gn = M.node(cx, cy, cz, a, iv);
          T          v  = rhs[gn * NT + i];
          if (ax)
            v -= ax[gn * NT + i];
          xs[l] = v;
        }
      __syncthreads();
      const int part = tid / G, lane = tid - part * G;
      const int a0 = part * CH, a1 = (a0 + CH < NL) ? a0 + CH : NL;
      // partial y_{k,i} = sum_{a in slice} S[a][k] x_{a,i}
      if (lane < NL)
        {
          T y[NT];
#pragma unroll
          for (int i = 0; i < NT; ++i)
            y[i] = T(0);
#pragma unroll 4
          for (int a = a0; a < a1; ++a)
            {
              const T s = ss[a * LD + lane];
#pragma unroll
              for (int i = 0; i < NT; ++i)
                y[i] += s * xs[a * NT + i];
            }
#pragma unroll
          for (int i = 0; i < NT; ++i)
            red[(part * NT + i) * NL + lane] = y[i];
        }
      __syncthreads();
      // z_{k,:} = W_k sum_parts y_{k,:}
      if (tid < NL)
        {
          const int k = tid;
          T         y[NT];
#pragma unroll
          for (int i = 0; i < NT; ++i)
            {
              T s = T(0);
#pragma unroll
              for (int q = 0; q < P; ++q)
                s += red[(q * NT + i) * NL + k];
              y[i] = s;
            }
          T z[NT];
          solve_time_block<T, NT>(lk, ta, tb, y, z);
#pragma unroll
          for (int i = 0; i < NT; ++i)
            ys[k * NT + i] = z[i];
        }
      __syncthreads();
      // partial out_{a,i} = sum_{k in slice} S[a][k] z_{k,i}
      if (lane < NL)
        {
          T o[NT];
#pragma unroll
          for (int i = 0; i < NT; ++i)
            o[i] = T(0);
#pragma unroll 4
          for (int k = a0; k < a1; ++k)
            {
              const T s = ss[lane * LD + k];
#pragma unroll
              for (int i = 0; i < NT; ++i)
                o[i] += s * ys[k * NT + i];
            }
#pragma unroll
          for (int i = 0; i < NT; ++i)
            red[(part * NT + i) * NL + lane] = o[i];
        }
      __syncthreads();
      using Sh = V2Shape<NL>;
      constexpr int n = Sh::n;
      for (int l = tid; l < NT * NL; l += 256)
        {
          const int i = l / NL, a = l - i * NL;
          T         s = T(0);
#pragma unroll
          for (int q = 0; q < P; ++q)
            s += red[(q * NT + i) * NL + a];
          float      iv;
          const long gn  = M.node(cx, cy, cz, a, iv);
          const T    val = omega * T(iv) * s;
          if constexpr (TWO_PASS)
            {
              const int la = a % n, lb = (a / n) % n, lc = (Sh::dim == 3) ? a / (n * n) : 0;
              cells[cell_slot<Sh::dim, n>(M.nc0, M.nc1, NT, cx, cy, cz, la, lb, lc, i)] = val;
            }
          else
            atomicAdd(&u[gn * NT + i], val);
        }
    }

    // Register-tiled factored apply (FP32, NL <= 128), 256 threads per
    // workgroup.  The threads form a 16 x 16 grid (ta, tk); thread (ta, tk)
    // keeps S[a][k] for rows a = ta + 16m (m < R) and the 16-byte column
    // groups g = tk + 16h (h < H) in registers, loaded with 16-byte loads
    // from the row-padded layout (16 lanes read 256 contiguous bytes of a
    // row); S takes no LDS, so several cells share a CU.  !PIPE: one cell
    // per workgroup.  PIPE: a persistent workgroup walks a contiguous range
    // of cells and loads the next cell's S tile, residual values and lambda_k
    // into a second register set while the current cell is processed (twice
    // the registers, one workgroup per CU).  Partial sums over the 16 tiles of a column
    // (of a row) meet in LDS laid out [tile][i][index], which makes the
    // reduction reads consecutive across lanes.  Four barriers per cell.
    // The corrections go to the cell buffer in the cell-block layout
    // (cell c's values contiguous from c * CS on, 128-byte lines; two-pass
    // form), which the node gather reads with the block stride CS.
    template <typename T, int NL, int NT>
    struct V3
    {
      static constexpr int    VW  = 16 / int(sizeof(T));
      static constexpr int    LDG = (NL + VW - 1) / VW * VW;              // global row stride
      static constexpr int    R   = (NL + 15) / 16;                        // rows per thread
      static constexpr int    H   = (NL + 16 * VW - 1) / (16 * VW);        // column groups per thread
      static constexpr int    KP  = 16 * VW * H;                           // padded column count
      static constexpr int    AP  = 16 * R;                                // padded row count
      static constexpr int    NX  = (NT * NL + 255) / 256;                 // residual values per thread
      static constexpr int    RS  = NT * AP + 1; // odd tile stride of the row partials
      static constexpr long   CS  = vanka_cell_block(NL, NT, sizeof(T)); // cell-block stride
      static constexpr size_t lds = sizeof(T) * (size_t(NT * NL) + size_t(NT * KP) +
                                                 (size_t(16) * NT * KP > size_t(16) * RS ?
                                                    size_t(16) * NT * KP : size_t(16) * RS));
    };

    // AXC: A x at the patch nodes is summed here from the operator's
    // cell-major buffer opc (the value its node gather would store), which
    // saves the gather launch of the residual; ax is not read.
    // NTS: the cell's results are stored non-temporally (cell buffers well
    // above the last-level cache); otherwise with ordinary stores
    template <typename T, int NL, int NT, bool PIPE, bool AXC = false, bool NTS = true>
    __device__ __forceinline__ void vanka_v3_range(const PatchMap &M, long n_cells,
                                                   const T *__restrict__ S,
                                                   const T *__restrict__ lam,
                                                   const T *__restrict__ talpha,
                                                   const T *__restrict__ tbeta,
                                                   const T *__restrict__ rhs,
                                                   const T *__restrict__ ax, T *__restrict__ cells,
                                                   T omega, int tid, T *smem,
                                                   const GridDesc *gd = nullptr,
                                                   const T *__restrict__ opc = nullptr)
    {
      using C = V3<T, NL, NT>;
      constexpr int VW = C::VW, LDG = C::LDG, R = C::R, H = C::H, KP = C::KP, AP = C::AP,
                    NX = C::NX, RS = C::RS;
      static_assert(NL <= 128, "register tiles hold at most 8 rows");
      typedef T vec_t __attribute__((ext_vector_type(VW)));
      using Sh = V2Shape<NL>;
      constexpr int n = Sh::n;
      T *const xs  = smem;         // [a][i]
      T *const ys  = xs + NT * NL; // [i][k], row stride KP
      T *const red = ys + NT * KP; // [16][NT][KP], or [16][RS] for the row partials
      const int ta = tid >> 4, tk = tid & 15;

      struct Buf
      {
        vec_t s[R][H];
        T     x[NX];
        T     lk; // eigenvalue of mode k = tid
      };
      auto load = [&](Buf &b, long cell) {
        // S is read once per sweep: streaming (nontemporal) loads
        const T *Sg = S + (size_t)cell * NL * LDG;
#pragma unroll
        for (int m = 0; m < R; ++m)
#pragma unroll
          for (int h = 0; h < H; ++h)
            {
              const int a = ta + 16 * m, g = tk + 16 * h;
              if (a < NL && VW * g < LDG)
                b.s[m][h] = __builtin_nontemporal_load(
                  reinterpret_cast<const vec_t *>(Sg + a * LDG + VW * g));
              else
                b.s[m][h] = vec_t(T(0));
            }
        int cx, cy, cz;
        M.cell_xyz(cell, cx, cy, cz);
#pragma unroll
        for (int j = 0; j < NX; ++j)
          {
            const int l = tid + 256 * j;
            T         v = T(0);
            if (l < NT * NL)
              {
                const int  a = l / NT, i = l - a * NT;
                float      iv;
                const long gn = node_ct<NL>(M, cx, cy, cz, a, iv);
                v             = rhs[gn * NT + i];
                if constexpr (AXC)
                  {
                    const int la = a % n, lb = (a / n) % n, lc = (Sh::dim == 3) ? a / (n * n) : 0;
                    v -= node_sum<Sh::dim, n, T>(*gd, NT, opc, (long)cx * (n - 1) + la,
                                                 (long)cy * (n - 1) + lb, (long)cz * (n - 1) + lc,
                                                 i);
                  }
                else if (ax)
                  v -= ax[gn * NT + i];
              }
            b.x[j] = v;
          }
        b.lk = (tid < NL) ? lam[(size_t)cell * NL + tid] : T(0);
      };
      auto process = [&](const Buf &b, long cell) {
#pragma unroll
        for (int j = 0; j < NX; ++j)
          {
            const int l = tid + 256 * j;
            if (l < NT * NL)
              xs[l] = b.x[j];
          }
        __syncthreads();
        // partial y[k][i] over this thread's rows
        {
          T y[H][VW][NT];
#pragma unroll
          for (int h = 0; h < H; ++h)
#pragma unroll
            for (int e = 0; e < VW; ++e)
#pragma unroll
              for (int i = 0; i < NT; ++i)
                y[h][e][i] = T(0);
#pragma unroll
          for (int m = 0; m < R; ++m)
            {
              const int a  = ta + 16 * m;
              const int ac = a < NL ? a : NL - 1; // rows past NL hold zeros
              T         xv[NT];
#pragma unroll
              for (int i = 0; i < NT; ++i)
                xv[i] = xs[ac * NT + i];
#pragma unroll
              for (int h = 0; h < H; ++h)
#pragma unroll
                for (int e = 0; e < VW; ++e)
#pragma unroll
                  for (int i = 0; i < NT; ++i)
                    y[h][e][i] += b.s[m][h][e] * xv[i];
            }
#pragma unroll
          for (int h = 0; h < H; ++h)
#pragma unroll
            for (int i = 0; i < NT; ++i)
              {
                vec_t v;
#pragma unroll
                for (int e = 0; e < VW; ++e)
                  v[e] = y[h][e][i];
                *reinterpret_cast<vec_t *>(red + (ta * NT + i) * KP + VW * (tk + 16 * h)) = v;
              }
        }
        __syncthreads();
        // z_k = W_k sum_ta y_k
        if (tid < NL)
          {
            const int k = tid;
            T         y[NT];
#pragma unroll
            for (int i = 0; i < NT; ++i)
              {
                T acc = T(0);
#pragma unroll
                for (int q = 0; q < 16; ++q)
                  acc += red[(q * NT + i) * KP + k];
                y[i] = acc;
              }
            T z[NT];
            solve_time_block<T, NT>(b.lk, talpha, tbeta, y, z);
#pragma unroll
            for (int i = 0; i < NT; ++i)
              ys[i * KP + k] = z[i];
          }
        else if (tid < KP)
#pragma unroll
          for (int i = 0; i < NT; ++i)
            ys[i * KP + tid] = T(0); // padding columns
        __syncthreads();
        // partial out[a][i] over this thread's columns
        {
          T o[R][NT];
#pragma unroll
          for (int m = 0; m < R; ++m)
#pragma unroll
            for (int i = 0; i < NT; ++i)
              o[m][i] = T(0);
#pragma unroll
          for (int h = 0; h < H; ++h)
            {
              vec_t zv[NT]; // z of this thread's VW columns of group h
#pragma unroll
              for (int i = 0; i < NT; ++i)
                zv[i] = *reinterpret_cast<const vec_t *>(ys + i * KP + VW * (tk + 16 * h));
#pragma unroll
              for (int e = 0; e < VW; ++e)
#pragma unroll
                for (int m = 0; m < R; ++m)
#pragma unroll
                  for (int i = 0; i < NT; ++i)
                    o[m][i] += b.s[m][h][e] * zv[i][e];
            }
#pragma unroll
          for (int m = 0; m < R; ++m)
#pragma unroll
            for (int i = 0; i < NT; ++i)
              red[tk * RS + i * AP + ta + 16 * m] = o[m][i];
        }
        __syncthreads();
        // node values of more than 16 bytes (FP64) with NTS: the cell's block
        // is staged in xs (free since the first product) and written by all
        // threads as consecutive 16-byte non-temporal stores (whole lines)
        constexpr bool STAGED = NTS && NT * sizeof(T) > 16 && (NT * sizeof(T)) % 16 == 0;
        if (tid < NL)
          {
            int cx, cy, cz;
            M.cell_xyz(cell, cx, cy, cz);
            const int a = tid;
            float     iv;
            (void)node_ct<NL>(M, cx, cy, cz, a, iv);
            const T sc = omega * T(iv);
            T       o[NT];
#pragma unroll
            for (int i = 0; i < NT; ++i)
              {
                T acc = T(0);
#pragma unroll
                for (int q = 0; q < 16; ++q)
                  acc += red[q * RS + i * AP + a];
                o[i] = sc * acc;
              }
            // the cell's block of the cell-block layout: node a's NT values
            // at a * NT; with NTS one 16-byte value per node is stored
            // non-temporally (the lanes of a wavefront write whole lines)
            T *const dst = cells + (size_t)cell * C::CS + (size_t)a * NT;
            if constexpr (STAGED)
              store_comps<T, NT>(xs + a * NT, o);
            else if constexpr (NTS && NT * sizeof(T) == 16)
              {
                vec_t v;
#pragma unroll
                for (int e = 0; e < VW; ++e)
                  v[e] = o[e];
                __builtin_nontemporal_store(v, reinterpret_cast<vec_t *>(dst));
              }
            else
              store_comps<T, NT>(dst, o);
          }
        if constexpr (STAGED)
          {
            __syncthreads();
            constexpr int NCH = NL * NT / VW; // 16-byte chunks of the block
            T *const      blk = cells + (size_t)cell * C::CS;
            for (int t = tid; t < NCH; t += 256)
              __builtin_nontemporal_store(*reinterpret_cast<const vec_t *>(xs + t * VW),
                                          reinterpret_cast<vec_t *>(blk) + t);
            if constexpr (PIPE)
              __syncthreads(); // the next cell's residual goes to xs
          }
      };

      if constexpr (!PIPE)
        {
          // one cell per workgroup, consecutive cells on one XCD
          const long cell = xcd_block((long)blockIdx.x, (long)gridDim.x);
          Buf        A;
          load(A, cell);
          process(A, cell);
        }
      else
        {
          const long per = (n_cells + gridDim.x - 1) / gridDim.x;
          const long c0  = (long)blockIdx.x * per;
          const long c1  = (c0 + per < n_cells) ? c0 + per : n_cells;
          if (c0 >= c1)
            return;
          Buf A, B;
          load(A, c0);
          for (long c = c0; c < c1; c += 2)
            {
              load(B, (c + 1 < c1) ? c + 1 : c1 - 1);
              process(A, c);
              if (c + 1 >= c1)
                break;
              load(A, (c + 2 < c1) ? c + 2 : c1 - 1);
              process(B, c + 1);
            }
        }
    }
  } // namespace
} // namespace stfem
