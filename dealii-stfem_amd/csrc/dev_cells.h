// Cell-major buffers of the two-pass (atomic-free) scatter (gfx950): the
// per-cell results of a cell kernel are stored in a fixed layout, and one
// thread per node sums the contributions of the cells sharing it in a fixed
// order.  Used by the matrix-free operator and by the Vanka smoother.
#pragma once

#include "kernels.h"

namespace stfem
{
  namespace
  {
    // gfx950 hands workgroup b of a launch to XCD b % 8 (round robin over
    // the 8 XCDs, each with its own L2).  Renumber the workgroups so that
    // XCD x works through one contiguous eighth of [0, G): neighbouring
    // cells, which share nodes of the input vectors and cache lines of the
    // cell-major buffer, then meet in one L2 instead of in eight.  A
    // bijection of [0, G) for every G.
    __device__ __forceinline__ long xcd_block(long b, long G)
    {
      constexpr long X = 8;
      const long     q = G / X, r = G % X, x = b % X;
      return x * q + (x < r ? x : r) + b / X;
    }

    // Position of the cell-local value (cell (cx,cy,cz), local node
    // (a,b,c), output component j) in the cell-major result buffer: for each
    // row of cells along x and local (c,b), the values of all cells of the
    // row lie contiguous in x ((cx, a) with a fastest), the nto components
    // of a (cell, node) innermost, so that the per-cell stores are coalesced
    // and a node gather reads each contribution's components as one vector.
    // In 2D the local line index is b (y) and c is unused.
    template <int DIM, int N>
    __host__ __device__ __forceinline__ long cell_slot(int nc0, int nc1, int nto, int cx, int cy,
                                                       int cz, int a, int b, int c, int j)
    {
      const long L = long(nc0) * N;
      if constexpr (DIM == 3)
        {
          const long row = long(cz) * nc1 + cy;
          return (((row * N + c) * N + b) * L + long(cx) * N + a) * nto + j;
        }
      else
        return ((long(cy) * N + b) * L + long(cx) * N + a) * nto + j;
    }
    template <int DIM, int N>
    __device__ __forceinline__ long cell_slot(const GridDesc &g, int nto, int cx, int cy, int cz,
                                              int a, int b, int c, int j)
    {
      return cell_slot<DIM, N>(g.nc[0], g.nc[1], nto, cx, cy, cz, a, b, c, j);
    }

    // Cell-block layout (the register-tiled Vanka sweep's results): the
    // n_loc * nto values of cell c (local node index a + N b + N^2 c, the
    // nto components innermost) are contiguous from c * cs on, with cs
    // rounded up to whole 128-byte lines so that each cell writes its values
    // as full lines (non-temporal stores that need no merging in L2).
    template <typename T>
    __host__ __device__ constexpr long cell_block_stride(int nloc, int nto)
    {
      return vanka_cell_block(nloc, nto, sizeof(T));
    }

    // NT contiguous values at p (p a multiple of NT values from a 256-byte
    // aligned base): 16-byte vector accesses when NT * sizeof(T) is a
    // multiple of 16, one 8-byte access for two floats, scalars otherwise
    template <typename T, int NT>
    __device__ __forceinline__ void load_comps(const T *__restrict__ p, T (&v)[NT])
    {
      constexpr int B = NT * int(sizeof(T));
      if constexpr (B % 16 == 0)
        {
          constexpr int PER = 16 / int(sizeof(T)); // values per 16 bytes
#pragma unroll
          for (int h = 0; h < NT / PER; ++h)
            {
              if constexpr (sizeof(T) == 8)
                {
                  const double2 q = reinterpret_cast<const double2 *>(p)[h];
                  v[2 * h]        = q.x;
                  v[2 * h + 1]    = q.y;
                }
              else
                {
                  const float4 q = reinterpret_cast<const float4 *>(p)[h];
                  v[4 * h]       = q.x;
                  v[4 * h + 1]   = q.y;
                  v[4 * h + 2]   = q.z;
                  v[4 * h + 3]   = q.w;
                }
            }
        }
      else if constexpr (B == 8 && sizeof(T) == 4)
        {
          const float2 q = *reinterpret_cast<const float2 *>(p);
          v[0]           = q.x;
          v[1]           = q.y;
        }
      else
#pragma unroll
        for (int j = 0; j < NT; ++j)
          v[j] = p[j];
    }
    template <typename T, int NT>
    __device__ __forceinline__ void store_comps(T *__restrict__ p, const T (&v)[NT])
    {
      constexpr int B = NT * int(sizeof(T));
      if constexpr (B % 16 == 0)
        {
          constexpr int PER = 16 / int(sizeof(T));
#pragma unroll
          for (int h = 0; h < NT / PER; ++h)
            {
              if constexpr (sizeof(T) == 8)
                reinterpret_cast<double2 *>(p)[h] = make_double2(v[2 * h], v[2 * h + 1]);
              else
                reinterpret_cast<float4 *>(p)[h] =
                  make_float4(v[4 * h], v[4 * h + 1], v[4 * h + 2], v[4 * h + 3]);
            }
        }
      else if constexpr (B == 8 && sizeof(T) == 4)
        *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
      else
#pragma unroll
        for (int j = 0; j < NT; ++j)
          p[j] = v[j];
    }

    // The (one to 2^dim) positions of node (gx, gy, gz)'s cell-local
    // values, component 0, in the order of summation of gather_node (cells
    // by increasing x, then y, then z); returns how many there are.
    // gather_node keeps its own copy of this logic: built on this function
    // the FP64 gather ran at half its speed.
    template <int DIM, int N>
    __device__ __forceinline__ int node_slots(const GridDesc &g, int nt, long gx, long gy, long gz,
                                              long slot[1 << DIM])
    {
      constexpr int p = N - 1;
      // the cells sharing the node and its local index in each, per direction
      int        cs[3][2] = {}, ls[3][2] = {}, nsh[3] = {1, 1, 1};
      const long gc[3] = {gx, gy, gz};
#pragma unroll
      for (int e = 0; e < DIM; ++e)
        {
          int c = int(gc[e] / p), l = int(gc[e] - long(c) * p);
          if (c == g.nc[e])
            c = g.nc[e] - 1, l = p;
          // lower neighbour first: fixed summation order
          if (l == 0 && c > 0)
            {
              cs[e][0] = c - 1;
              ls[e][0] = p;
              cs[e][1] = c;
              ls[e][1] = 0;
              nsh[e]   = 2;
            }
          else
            {
              cs[e][0] = c;
              ls[e][0] = l;
            }
        }
      int ns = 0;
#pragma unroll
      for (int kz = 0; kz < (DIM == 3 ? 2 : 1); ++kz)
#pragma unroll
        for (int ky = 0; ky < 2; ++ky)
#pragma unroll
          for (int kx = 0; kx < 2; ++kx)
            {
              const bool use = kx < nsh[0] && ky < nsh[1] && (DIM == 2 || kz < nsh[2]);
              if (use)
                slot[ns] = (DIM == 3) ? cell_slot<DIM, N>(g, nt, cs[0][kx], cs[1][ky], cs[2][kz],
                                                          ls[0][kx], ls[1][ky], ls[2][kz], 0)
                                      : cell_slot<DIM, N>(g, nt, cs[0][kx], cs[1][ky], 0,
                                                          ls[0][kx], ls[1][ky], 0, 0);
              ns += use ? 1 : 0;
            }
      return ns;
    }

    // Whether node (gx, gy, gz) lies on a Dirichlet face.
    template <int DIM>
    __device__ __forceinline__ bool dirichlet_node(const GridDesc &g, long gx, long gy, long gz)
    {
      const int d = g.dirichlet;
      bool bnd = (gx == 0 && (d & 1)) || (gx == g.nn[0] - 1 && (d & 2)) || (gy == 0 && (d & 4)) ||
                 (gy == g.nn[1] - 1 && (d & 8));
      if (DIM == 3)
        bnd = bnd || (gz == 0 && (d & 16)) || (gz == g.nn[2] - 1 && (d & 32));
      return bnd;
    }

    // Component j of node (gx, gy, gz) summed from a cell-major buffer of
    // nto components, zero on Dirichlet nodes: the value gather_node stores
    // in mode 0 (same order of summation, so the same bits), computed where
    // it is consumed.
    template <int DIM, int N, typename T>
    __device__ __forceinline__ T node_sum(const GridDesc &g, int nto, const T *__restrict__ cells,
                                          long gx, long gy, long gz, int j)
    {
      if (dirichlet_node<DIM>(g, gx, gy, gz))
        return T(0);
      long      slot[1 << DIM];
      const int ns = node_slots<DIM, N>(g, nto, gx, gy, gz, slot);
      T s = T(0);
      for (int k = 0; k < ns; ++k)
        s += cells[slot[k] + j];
      return s;
    }

    // Second pass of the two-pass forms: node `node` of the level sums its
    // (one to 2^dim) cell-local results from the cell-major buffer, in a
    // fixed order (cells by increasing x, then y, then z), and stores
    //   mode 0: dst = sum,  mode 1: dst += sum,  mode 2: dst = rhs - sum.
    // SKIP_DIR (operator): Dirichlet nodes get zero (modes 0, 2) or stay
    // unchanged (mode 1); otherwise (Vanka, whose patches cover them) they
    // are summed like every other node.
    // NT > 0: compile-time number of components (all loads issued before
    // the sums); NT == 0: nto at run time.  CB: the buffer is in the
    // cell-block layout of block stride cs (same contributions, same order).
    template <int DIM, int N, int NT, typename T, bool SKIP_DIR = true, bool CB = false>
    __device__ __forceinline__ void gather_node(const GridDesc &g, int nto,
                                                const T *__restrict__ cells, T *__restrict__ dst,
                                                const T *__restrict__ rhs, int mode, long node,
                                                long cs = 0)
    {
      constexpr int p   = N - 1;
      const long    nn0 = g.nn[0], nn1 = g.nn[1];
      // 32-bit splits (levels have fewer than 2^31 nodes)
      const unsigned un  = unsigned(node), u0 = unsigned(nn0), u1 = unsigned(nn1);
      const unsigned ur  = un / u0;
      const long     gx  = long(un - ur * u0);
      const long     gy  = (DIM == 3) ? long(ur % u1) : long(ur);
      const long     gz  = (DIM == 3) ? long(ur / u1) : 0;
      const int     d   = g.dirichlet;
      bool bnd = (gx == 0 && (d & 1)) || (gx == nn0 - 1 && (d & 2)) || (gy == 0 && (d & 4)) ||
                 (gy == nn1 - 1 && (d & 8));
      if (DIM == 3)
        bnd = bnd || (gz == 0 && (d & 16)) || (gz == g.nn[2] - 1 && (d & 32));
      if (!SKIP_DIR)
        bnd = false;
      const int nt  = NT > 0 ? NT : nto;
      T *const  out = dst + node * nt;
      if (bnd)
        {
          if (mode != 1)
            for (int j = 0; j < nt; ++j)
              out[j] = T(0);
          return;
        }
      // per direction the first cell holding the node and the node's local
      // index there, and whether the next cell shares it (lower neighbour
      // first: fixed summation order)
      int        c0[3] = {}, l0[3] = {}, nsh[3] = {1, 1, 1};
      const long gc[3] = {gx, gy, gz};
#pragma unroll
      for (int e = 0; e < DIM; ++e)
        {
          int c = int(unsigned(gc[e]) / unsigned(p)), l = int(gc[e]) - c * p;
          if (c == g.nc[e])
            c = g.nc[e] - 1, l = p;
          if (l == 0 && c > 0)
            {
              c0[e]  = c - 1;
              l0[e]  = p;
              nsh[e] = 2;
            }
          else
            {
              c0[e] = c;
              l0[e] = l;
            }
        }
      // the next cell's copy of a shared node: along x the adjacent entry of
      // the same row ((cx, p) -> (cx + 1, 0)); along y and z the rows of the
      // next cell with local index 0 instead of p.  Cell blocks: the next
      // cell's block with the local index lowered by p (x), p N (y), p N^2 (z).
      constexpr int MAXS = 1 << DIM;
      long          base, dx, dy, dz;
      if constexpr (CB)
        {
          const long cell = (DIM == 3) ? c0[0] + long(g.nc[0]) * (c0[1] + long(g.nc[1]) * c0[2])
                                       : c0[0] + long(g.nc[0]) * c0[1];
          const int  loc  = (DIM == 3) ? l0[0] + N * (l0[1] + N * l0[2]) : l0[0] + N * l0[1];
          base            = cell * cs + long(loc) * nt;
          dx              = cs - long(p) * nt;
          dy              = long(g.nc[0]) * cs - long(p) * N * nt;
          dz = (DIM == 3) ? long(g.nc[0]) * g.nc[1] * cs - long(p) * N * N * nt : 0;
        }
      else
        {
          const long L = long(g.nc[0]) * N;
          base = (DIM == 3) ? cell_slot<DIM, N>(g, nt, c0[0], c0[1], c0[2], l0[0], l0[1], l0[2], 0)
                            : cell_slot<DIM, N>(g, nt, c0[0], c0[1], 0, l0[0], l0[1], 0, 0);
          dx   = nt;
          dy   = (DIM == 3) ? (long(N) * N - p) * L * nt : (long(N) - p) * L * nt;
          dz   = (DIM == 3) ? (long(g.nc[1]) * N * N - long(p) * N) * L * nt : 0;
        }
      // contribution k = kx + 2 ky + 4 kz is used where the node is shared
      // in those directions; unused ones add exact zeros, so the sum is the
      // one over the used contributions in the fixed order
      auto use = [&](int k) {
        return (k & 1) < nsh[0] && ((k >> 1) & 1) < nsh[1] && (DIM == 2 || (k >> 2) < nsh[2]);
      };
      auto off = [&](int k) { return (k & 1) * dx + ((k >> 1) & 1) * dy + (k >> 2) * dz; };
      if constexpr (NT > 0)
        {
          // each contribution's NT components are contiguous (the slot is a
          // multiple of NT): vector loads where NT * sizeof(T) is 8 or a
          // multiple of 16 bytes, and the same for the node's values
          T v[MAXS][NT];
#pragma unroll
          for (int k = 0; k < MAXS; ++k)
            {
              if (use(k))
                load_comps<T, NT>(cells + base + off(k), v[k]);
              else
#pragma unroll
                for (int j = 0; j < NT; ++j)
                  v[k][j] = T(0);
            }
          T o[NT], r[NT];
          if (mode == 1)
            load_comps<T, NT>(out, o);
          if (mode == 2)
            load_comps<T, NT>(rhs + node * NT, r);
#pragma unroll
          for (int j = 0; j < NT; ++j)
            {
              T s = v[0][j];
#pragma unroll
              for (int k = 1; k < MAXS; ++k)
                s += v[k][j];
              o[j] = mode == 0 ? s : mode == 1 ? o[j] + s : r[j] - s;
            }
          store_comps<T, NT>(out, o);
        }
      else
        for (int j = 0; j < nt; ++j)
          {
            T s = T(0);
            for (int k = 0; k < MAXS; ++k)
              if (use(k))
                s += cells[base + off(k) + j];
            out[j] = mode == 0 ? s : mode == 1 ? out[j] + s : rhs[node * nt + j] - s;
          }
    }
  } // namespace
} // namespace stfem
