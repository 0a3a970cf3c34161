// Device code of the multigrid transfers (gfx950), shared by the
// one-launch-per-transfer kernels and the coarse-level kernel.
//
// Time (tau / k) transfers: MGTwoLevelTransferST::prolongate_and_add /
// restrict_and_add (tests/include/gmg.h:121-135 via tensorproduct_add,
// operators.h:47-58 of nlsmrg/dealii-stfem) apply a small dense matrix to
// the nt time values of every node.
//
// Space (h) transfers: the embedding of Q_p on the coarse grid into Q_p on
// the once-refined grid (deal.II MGTwoLevelTransfer as built at gmg.h:371-380)
// is, on a structured grid, a Kronecker product of banded 1D operators with
// zero-Dirichlet rows/columns removed, applied as one 1D sweep per direction.
#pragma once

#include "kernels.h"

namespace stfem
{
  namespace
  {
    // entry idx = node*ntd + a of dst: (=|+=) sum_b M(a,b) src[node][b]
    template <typename T>
    __device__ __forceinline__ void time_transfer_item(long idx, const T *__restrict__ M, int ntd,
                                                       int nts, const T *__restrict__ src,
                                                       T *__restrict__ dst, int add)
    {
      const long node = idx / ntd;
      const int  a    = int(idx - node * ntd);
      T          s    = T(0);
      for (int b = 0; b < nts; ++b)
        s += M[a * nts + b] * src[node * nts + b];
      dst[idx] = add ? dst[idx] + s : s;
    }

    // entry idx of the output of a 1D banded operator along direction d:
    // output dims are n with n[d] replaced by m, nt components interleaved
    template <typename T>
    __device__ __forceinline__ void sweep_item(long idx, const T *__restrict__ in,
                                               T *__restrict__ out, int nt, int d, int n0, int n1,
                                               int m, const int *__restrict__ cnt,
                                               const int *__restrict__ col,
                                               const T *__restrict__ w, int K, int add)
    {
      const long o0 = (d == 0) ? m : n0, o1 = (d == 1) ? m : n1;
      const long pt   = idx / nt;
      const int  comp = int(idx - pt * nt);
      const long i0 = pt % o0, r = pt / o0, i1 = r % o1, i2 = r / o1;
      const long od = (d == 0) ? i0 : (d == 1) ? i1 : i2;
      const long stride = (d == 0) ? 1 : (d == 1) ? (long)n0 : (long)n0 * n1;
      // input index with the d coordinate set to 0
      const long j0 = (d == 0) ? 0 : i0, j1 = (d == 1) ? 0 : i1, j2 = (d == 2) ? 0 : i2;
      const long ib = j0 + (long)n0 * (j1 + (long)n1 * j2);
      const int  c  = cnt[od];
      T          s  = T(0);
      for (int k = 0; k < c; ++k)
        s += w[od * K + k] * in[(ib + (long)col[od * K + k] * stride) * nt + comp];
      out[idx] = add ? out[idx] + s : s;
    }

    inline long sweep_total(int nt, int d, const int n_in[3], int m_out)
    {
      long o[3] = {n_in[0], n_in[1], n_in[2]};
      o[d]      = m_out;
      return o[0] * o[1] * o[2] * nt;
    }
  } // namespace
} // namespace stfem
