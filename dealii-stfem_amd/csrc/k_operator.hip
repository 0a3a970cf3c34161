// Fused matrix-free space-time operator for gfx950, one launch per
// application.
//
// Replaces, in one pass over the cells, the 2*nt MatrixFreeOperator::vmult
// calls and the nt^2 block axpys of SystemMatrix::vmult
// (tests/include/operators.h:94-122 and :487-548 of nlsmrg/dealii-stfem).
// Work decomposition: a 256-thread workgroup holds CPB cells; the per-cell
// device code (sum factorization, on-the-fly geometry) is in dev_operator.h.
// Two-pass form (default): the cell kernel stores each cell's results into a
// cell-major buffer with coalesced plain stores and a node kernel sums the
// (up to 2^dim) contributions of every node in a fixed order.  This avoids
// the device-scope float atomics of the one-pass form, which are performed
// one at a time at the memory side and cost more than the whole cell pass.
#include "dev_operator_wave.h"

namespace stfem
{
  namespace
  {
    template <int DIM, int N, int JB, typename T, bool TO_CELLS>
    __global__ void __launch_bounds__(256)
      k_st_apply(const StOpArgs<T> A, const T *__restrict__ src, T *__restrict__ dst, int jb)
    {
      using Ly = Layout<DIM, N>;
      __shared__ T lds0[Ly::CPB * Ly::NC];
      __shared__ T lds1[Ly::CPB * Ly::NC];
      __shared__ T vlds[Ly::CPB * (1 << DIM) * DIM];
      // output components [j0, jn) of this grid row
      const int j0 = (int)blockIdx.y * jb;
      const int jn = (j0 + jb < A.nt_out) ? j0 + jb : A.nt_out;
      st_apply_cells<DIM, N, JB, T, TO_CELLS>(A, src, dst, j0, jn,
                                              xcd_block((long)blockIdx.x, (long)gridDim.x),
                                              (int)threadIdx.x, lds0, lds1, vlds);
    }

    template <int N, int NTI, typename T>
    __global__ void __launch_bounds__(WaveLayout<N>::BLOCK)
      k_st_apply_wave(const StOpArgs<T> A, const T *__restrict__ src, T *__restrict__ cells, int jb)
    {
      using W = WaveLayout<N>;
      __shared__ T lds[W::GPB * W::CPG * W::PER];
      const int j0 = (int)blockIdx.y * jb;
      const int jn = (j0 + jb < A.nt_out) ? j0 + jb : A.nt_out;
      st_apply_wave<N, NTI, T>(A, src, cells, j0, jn, xcd_block((long)blockIdx.x, (long)gridDim.x),
                               lds);
    }

    template <int DIM, int N, int NT, typename T, bool SKIP_DIR, bool CB>
    __global__ void __launch_bounds__(256)
      k_gather_nodes(const GridDesc g, int nto, const T *__restrict__ cells, T *__restrict__ dst,
                     const T *__restrict__ rhs, int mode, long cs)
    {
      const long node = xcd_block((long)blockIdx.x, (long)gridDim.x) * 256 + threadIdx.x;
      if (node < g.n_nodes)
        gather_node<DIM, N, NT, T, SKIP_DIR, CB>(g, nto, cells, dst, rhs, mode, node, cs);
    }

    // cs > 0: the buffer is in the cell-block layout of stride cs
    template <int DIM, int N, typename T, bool SKIP_DIR, bool CB = false>
    void gather_dn(const GridDesc &g, int nto, const T *cells, T *dst, const T *rhs, int mode,
                   hipStream_t s, long cs = 0)
    {
      if constexpr (!CB)
        if (cs > 0)
          return gather_dn<DIM, N, T, SKIP_DIR, true>(g, nto, cells, dst, rhs, mode, s, cs);
      const dim3 grid{(unsigned)((g.n_nodes + 255) / 256), 1, 1};
      // the compile-time forms access a node's components as vectors: node
      // vectors not 16-byte aligned take the run-time form
      const bool aligned = reinterpret_cast<uintptr_t>(dst) % 16 == 0 &&
                           reinterpret_cast<uintptr_t>(rhs) % 16 == 0;
      switch (aligned ? nto : 0)
        {
          case 1: k_gather_nodes<DIM, N, 1, T, SKIP_DIR, CB><<<grid, 256, 0, s>>>(g, 1, cells, dst, rhs, mode, cs); break;
          case 2: k_gather_nodes<DIM, N, 2, T, SKIP_DIR, CB><<<grid, 256, 0, s>>>(g, 2, cells, dst, rhs, mode, cs); break;
          case 3: k_gather_nodes<DIM, N, 3, T, SKIP_DIR, CB><<<grid, 256, 0, s>>>(g, 3, cells, dst, rhs, mode, cs); break;
          case 4: k_gather_nodes<DIM, N, 4, T, SKIP_DIR, CB><<<grid, 256, 0, s>>>(g, 4, cells, dst, rhs, mode, cs); break;
          default:
            k_gather_nodes<DIM, N, 0, T, SKIP_DIR, CB><<<grid, 256, 0, s>>>(g, nto, cells, dst, rhs, mode, cs);
            break;
        }
      STFEM_HIP(hipGetLastError());
    }
    template <typename T, bool SKIP_DIR>
    void gather_any(const GridDesc &g, int nto, const T *cells, T *dst, const T *rhs, int mode,
                    hipStream_t s, long cs = 0)
    {
      const int N = g.p + 1;
#define STFEM_GATHER_CASE(D, NN)                                                                   \
  if (g.dim == D && N == NN)                                                                       \
    return gather_dn<D, NN, T, SKIP_DIR>(g, nto, cells, dst, rhs, mode, s, cs);
      STFEM_GATHER_CASE(2, 2) STFEM_GATHER_CASE(2, 3) STFEM_GATHER_CASE(2, 4)
      STFEM_GATHER_CASE(2, 5) STFEM_GATHER_CASE(2, 6) STFEM_GATHER_CASE(3, 2)
      STFEM_GATHER_CASE(3, 3) STFEM_GATHER_CASE(3, 4) STFEM_GATHER_CASE(3, 5)
      STFEM_GATHER_CASE(3, 6)
#undef STFEM_GATHER_CASE
      throw Error(STATUS_ERR_UNSUPPORTED, "gather: unsupported dim/degree");
    }

    template <int DIM, int N, typename T, bool TO_CELLS>
    void dispatch_jb(const StOpArgs<T> &a, const T *src, T *dst, hipStream_t s)
    {
      using Ly              = Layout<DIM, N>;
      const int  nto        = a.nt_out;
      if constexpr (DIM == 3 && TO_CELLS)
        if (a.kernel == 1 && a.nt_in <= 4)
          {
            // all output components in one pass over the inputs; on small
            // grids one component per grid row
            using W           = WaveLayout<N>;
            const long blocks = (a.g.n_cells + W::CPB - 1) / W::CPB;
            const int  jb     = blocks < a.split_below ? 1 : nto;
            const dim3 grid{(unsigned)blocks, (unsigned)((nto + jb - 1) / jb), 1};
            constexpr unsigned WB = W::BLOCK;
            switch (a.nt_in)
              {
                case 1: k_st_apply_wave<N, 1, T><<<grid, WB, 0, s>>>(a, src, dst, jb); break;
                case 2: k_st_apply_wave<N, 2, T><<<grid, WB, 0, s>>>(a, src, dst, jb); break;
                case 3: k_st_apply_wave<N, 3, T><<<grid, WB, 0, s>>>(a, src, dst, jb); break;
                default: k_st_apply_wave<N, 4, T><<<grid, WB, 0, s>>>(a, src, dst, jb); break;
              }
            STFEM_HIP(hipGetLastError());
            return;
          }
      const int  mj         = (a.max_jb >= 1 && a.max_jb <= 4) ? a.max_jb : 4;
      const long blocks     = (a.g.n_cells + Ly::CPB - 1) / Ly::CPB;
      int        chunks     = (nto + mj - 1) / mj;
      if (blocks < a.split_below)
        chunks = nto;
      const int  jb         = (nto + chunks - 1) / chunks;
      chunks                = (nto + jb - 1) / jb;
      // the chunks of output components run side by side as grid rows
      const dim3 grid{(unsigned)blocks, (unsigned)chunks, 1};
      const dim3 block{(unsigned)(Ly::CPB * Ly::TPC), 1, 1};
      switch (jb)
        {
          case 1: k_st_apply<DIM, N, 1, T, TO_CELLS><<<grid, block, 0, s>>>(a, src, dst, jb); break;
          case 2: k_st_apply<DIM, N, 2, T, TO_CELLS><<<grid, block, 0, s>>>(a, src, dst, jb); break;
          case 3: k_st_apply<DIM, N, 3, T, TO_CELLS><<<grid, block, 0, s>>>(a, src, dst, jb); break;
          default: k_st_apply<DIM, N, 4, T, TO_CELLS><<<grid, block, 0, s>>>(a, src, dst, jb); break;
        }
      STFEM_HIP(hipGetLastError());
    }
  } // namespace

  void upload_shape_tables(const double *S, const double *D, const double *W, const double *X,
                           int N)
  {
    STFEM_REQUIRE(N >= 2 && N <= MAX_N, "upload_shape_tables: unsupported degree");
    upload_tables_here(S, D, W, X, N);
  }

  namespace
  {
    // mode 0: dst = A src, 1: dst += A src, 2: dst = rhs - A src, 3: the
    // cell pass only
    template <int DIM, int N, typename T>
    void apply_dn(const StOpArgs<T> &a, const T *src, T *dst, T *cells, const T *rhs, int mode,
                  hipStream_t s)
    {
      const long m = a.g.n_nodes * a.nt_out;
      if (mode == 3)
        {
          STFEM_REQUIRE(cells, "st_apply: the cell pass alone needs the two-pass operator");
          return dispatch_jb<DIM, N, T, true>(a, src, cells, s);
        }
      if (!cells)
        {
          if (mode == 2)
            throw Error(STATUS_ERR_UNSUPPORTED, "st_apply: residual form needs the two-pass operator");
          if (mode == 0)
            launch_fill_zero<T>(dst, size_t(m), s);
          return dispatch_jb<DIM, N, T, false>(a, src, dst, s);
        }
      dispatch_jb<DIM, N, T, true>(a, src, cells, s);
      gather_dn<DIM, N, T, true>(a.g, a.nt_out, cells, dst, rhs, mode, s);
    }
  } // namespace

  template <typename T>
  void launch_st_apply(const StOpArgs<T> &a, const T *src, T *dst, T *cells, int mode,
                       const T *rhs, hipStream_t s)
  {
    STFEM_REQUIRE(a.nt_in >= 1 && a.nt_out >= 1, "st_apply: bad number of time components");
    if (a.g.n_cells == 0)
      return;
    const int N = a.g.p + 1;
    if (a.g.dim == 2)
      switch (N)
        {
          case 2: return apply_dn<2, 2, T>(a, src, dst, cells, rhs, mode, s);
          case 3: return apply_dn<2, 3, T>(a, src, dst, cells, rhs, mode, s);
          case 4: return apply_dn<2, 4, T>(a, src, dst, cells, rhs, mode, s);
          case 5: return apply_dn<2, 5, T>(a, src, dst, cells, rhs, mode, s);
          case 6: return apply_dn<2, 6, T>(a, src, dst, cells, rhs, mode, s);
          default: break;
        }
    else if (a.g.dim == 3)
      switch (N)
        {
          case 2: return apply_dn<3, 2, T>(a, src, dst, cells, rhs, mode, s);
          case 3: return apply_dn<3, 3, T>(a, src, dst, cells, rhs, mode, s);
          case 4: return apply_dn<3, 4, T>(a, src, dst, cells, rhs, mode, s);
          case 5: return apply_dn<3, 5, T>(a, src, dst, cells, rhs, mode, s);
          case 6: return apply_dn<3, 6, T>(a, src, dst, cells, rhs, mode, s);
          default: break;
        }
    throw Error(STATUS_ERR_UNSUPPORTED, "st_apply: unsupported dim/degree");
  }

  template <typename T>
  void launch_gather_cells(const GridDesc &g, int nto, const T *cells, T *dst, const T *rhs,
                           int mode, bool skip_dirichlet, hipStream_t s, long cs)
  {
    STFEM_REQUIRE(nto >= 1 && mode >= 0 && mode <= 2, "gather: bad arguments");
    STFEM_REQUIRE(mode != 2 || rhs, "gather: mode 2 needs rhs");
    STFEM_REQUIRE(cs == 0 || (cs >= long(nto) * (g.dim == 3 ? (g.p + 1) * (g.p + 1) * (g.p + 1)
                                                            : (g.p + 1) * (g.p + 1)) &&
                              cs % nto == 0),
                  "gather: cell-block stride below the cell's values");
    if (g.n_nodes == 0)
      return;
    if (skip_dirichlet)
      gather_any<T, true>(g, nto, cells, dst, rhs, mode, s, cs);
    else
      gather_any<T, false>(g, nto, cells, dst, rhs, mode, s, cs);
  }
  template void launch_gather_cells<double>(const GridDesc &, int, const double *, double *,
                                            const double *, int, bool, hipStream_t, long);
  template void launch_gather_cells<float>(const GridDesc &, int, const float *, float *,
                                           const float *, int, bool, hipStream_t, long);

  template void launch_st_apply<double>(const StOpArgs<double> &, const double *, double *,
                                        double *, int, const double *, hipStream_t);
  template void launch_st_apply<float>(const StOpArgs<float> &, const float *, float *, float *,
                                       int, const float *, hipStream_t);
} // namespace stfem
