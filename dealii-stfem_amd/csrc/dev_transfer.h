// Device code of the multigrid transfers (gfx950), used by the kernels of
// k_transfer.hip.
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

#include "dev_cells.h"

namespace stfem
{
  namespace
  {
    // entry idx = node*ntd + a of dst: (=|+=) sum_b M(a,b) src[node][b]; I is the
    // index type (32-bit unsigned whenever every index fits, which avoids the
    // emulated 64-bit divisions)
    template <typename T, typename I>
    __device__ __forceinline__ void time_transfer_item(I idx, const T *__restrict__ M, int ntd,
                                                       int nts, const T *__restrict__ src,
                                                       T *__restrict__ dst, int add)
    {
      const I   node = idx / I(ntd);
      const int a    = int(idx - node * I(ntd));
      T         s    = T(0);
#pragma unroll 4
      for (int b = 0; b < nts; ++b)
        s += M[a * nts + b] * src[node * I(nts) + I(b)];
      dst[idx] = add ? dst[idx] + s : s;
    }

    // entry idx of the output of a 1D banded operator along direction d:
    // output dims are n with n[d] replaced by m, nt components interleaved;
    // I as for time_transfer_item, covering the input and the output.  With
    // KM > 0 (row stride K <= KM) the count and the column and weight loads
    // of all K taps are issued together, then the input loads of the row's
    // cnt taps, so that a thread waits for two memory round trips instead of
    // one per tap (padded taps: column 0, weight 0, never summed); the taps
    // are summed in order as with KM = 0.  The KM = 0 loop is unrolled by
    // four for the same reason: the loads of four taps are then in flight
    // together.
    template <typename T, typename I, int KM>
    __device__ __forceinline__ void sweep_item(I idx, const T *__restrict__ in,
                                               T *__restrict__ out, int nt, int d, int n0, int n1,
                                               int m, const int *__restrict__ cnt,
                                               const int *__restrict__ col,
                                               const T *__restrict__ w, int K, int add)
    {
      const I   o0 = I((d == 0) ? m : n0), o1 = I((d == 1) ? m : n1);
      const I   pt   = idx / I(nt);
      const int comp = int(idx - pt * I(nt));
      const I   i0 = pt % o0, r = pt / o0, i1 = r % o1, i2 = r / o1;
      const int od = int((d == 0) ? i0 : (d == 1) ? i1 : i2);
      const I   stride = (d == 0) ? I(1) : (d == 1) ? I(n0) : I(n0) * I(n1);
      // input index with the d coordinate set to 0
      const I j0 = (d == 0) ? I(0) : i0, j1 = (d == 1) ? I(0) : i1, j2 = (d == 2) ? I(0) : i2;
      const I ib = j0 + I(n0) * (j1 + I(n1) * j2);
      const int c = cnt[od];
      T         s = T(0);
      if constexpr (KM == 0)
        {
#pragma unroll 4
          for (int k = 0; k < c; ++k)
            s += w[od * K + k] * in[(ib + I(col[od * K + k]) * stride) * I(nt) + I(comp)];
        }
      else
        {
          int cl[KM];
          T   wl[KM], xl[KM];
#pragma unroll
          for (int k = 0; k < KM; ++k)
            if (k < K)
              {
                cl[k] = col[od * K + k];
                wl[k] = w[od * K + k];
              }
#pragma unroll
          for (int k = 0; k < KM; ++k)
            if (k < c)
              xl[k] = in[(ib + I(cl[k]) * stride) * I(nt) + I(comp)];
#pragma unroll
          for (int k = 0; k < KM; ++k)
            if (k < c)
              s += wl[k] * xl[k];
        }
      out[idx] = add ? out[idx] + s : s;
    }

    // node pt of the output of the same 1D banded operator, all NT
    // components at once: the row's tables are read once and every tap loads
    // the node's NT contiguous values (16-byte accesses where NT * sizeof(T)
    // is a multiple of 16, the vectors 16-byte aligned); the sums run over
    // the taps in order, as in sweep_item, so the results are the same
    template <typename T, typename I, int KM, int NT>
    __device__ __forceinline__ void sweep_node(I pt, const T *__restrict__ in,
                                               T *__restrict__ out, int d, int n0, int n1, int m,
                                               const int *__restrict__ cnt,
                                               const int *__restrict__ col,
                                               const T *__restrict__ w, int K, int add)
    {
      const I   o0 = I((d == 0) ? m : n0), o1 = I((d == 1) ? m : n1);
      const I   i0 = pt % o0, r = pt / o0, i1 = r % o1, i2 = r / o1;
      const int od = int((d == 0) ? i0 : (d == 1) ? i1 : i2);
      const I   stride = (d == 0) ? I(1) : (d == 1) ? I(n0) : I(n0) * I(n1);
      const I j0 = (d == 0) ? I(0) : i0, j1 = (d == 1) ? I(0) : i1, j2 = (d == 2) ? I(0) : i2;
      const I ib = j0 + I(n0) * (j1 + I(n1) * j2);
      const int c = cnt[od];
      T         s[NT];
#pragma unroll
      for (int j = 0; j < NT; ++j)
        s[j] = T(0);
      auto tap = [&](T wk, int ck) {
        T x[NT];
        load_comps<T, NT>(in + (ib + I(ck) * stride) * I(NT), x);
#pragma unroll
        for (int j = 0; j < NT; ++j)
          s[j] += wk * x[j];
      };
      if constexpr (KM == 0)
        {
#pragma unroll 4
          for (int k = 0; k < c; ++k)
            tap(w[od * K + k], col[od * K + k]);
        }
      else
        {
          int cl[KM];
          T   wl[KM];
#pragma unroll
          for (int k = 0; k < KM; ++k)
            if (k < K)
              {
                cl[k] = col[od * K + k];
                wl[k] = w[od * K + k];
              }
          T xl[KM][NT];
#pragma unroll
          for (int k = 0; k < KM; ++k)
            if (k < c)
              load_comps<T, NT>(in + (ib + I(cl[k]) * stride) * I(NT), xl[k]);
#pragma unroll
          for (int k = 0; k < KM; ++k)
            if (k < c)
#pragma unroll
              for (int j = 0; j < NT; ++j)
                s[j] += wl[k] * xl[k][j];
        }
      T *const o = out + pt * I(NT);
      if (add)
        {
          T v[NT];
          load_comps<T, NT>(o, v);
#pragma unroll
          for (int j = 0; j < NT; ++j)
            s[j] = v[j] + s[j];
        }
      store_comps<T, NT>(o, s);
    }

    inline long sweep_total(int nt, int d, const int n_in[3], int m_out)
    {
      long o[3] = {n_in[0], n_in[1], n_in[2]};
      o[d]      = m_out;
      return o[0] * o[1] * o[2] * nt;
    }
  } // namespace
} // namespace stfem
