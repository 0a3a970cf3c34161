// Multigrid transfer kernels for gfx950.
//
// Time (tau / k) transfers: MGTwoLevelTransferST::prolongate_and_add /
// restrict_and_add (tests/include/gmg.h:121-135 via tensorproduct_add,
// operators.h:47-58 of nlsmrg/dealii-stfem) apply a small dense matrix to
// the nt time values of every node: with the node-major layout this is one
// streaming pass.
//
// Space (h) transfers: the embedding of Q_p on the coarse grid into Q_p on
// the once-refined grid (deal.II MGTwoLevelTransfer as built at gmg.h:371-380)
// is, on a structured grid, a Kronecker product of banded 1D operators with
// zero-Dirichlet rows/columns removed.  It is applied as one 1D sweep per
// direction; restriction uses the transposed 1D operators.
#include "kernels.h"

namespace stfem
{
  namespace
  {
    template <typename T>
    __global__ void k_time_transfer(long n_nodes, const T *__restrict__ M, int ntd, int nts,
                                    const T *__restrict__ src, T *__restrict__ dst, int add)
    {
      const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (idx >= n_nodes * ntd)
        return;
      const long node = idx / ntd;
      const int  a    = int(idx - node * ntd);
      T          s    = T(0);
      for (int b = 0; b < nts; ++b)
        s += M[a * nts + b] * src[node * nts + b];
      dst[idx] = add ? dst[idx] + s : s;
    }

    template <typename T>
    __global__ void k_sweep(const T *__restrict__ in, T *__restrict__ out, int nt, int d, int n0,
                            int n1, int n2, int m, const int *__restrict__ cnt,
                            const int *__restrict__ col, const T *__restrict__ w, int K, int add)
    {
      // output dims: n with n[d] replaced by m
      const long o0 = (d == 0) ? m : n0, o1 = (d == 1) ? m : n1, o2 = (d == 2) ? m : n2;
      const long total = o0 * o1 * o2 * nt;
      const long idx   = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (idx >= total)
        return;
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
  } // namespace

  template <typename T>
  void launch_time_transfer(long n_nodes, const T *M, int ntd, int nts, const T *src, T *dst,
                            bool add, hipStream_t s)
  {
    const long total = n_nodes * ntd;
    if (total == 0)
      return;
    k_time_transfer<T><<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s>>>(
      n_nodes, M, ntd, nts, src, dst, add ? 1 : 0);
    STFEM_HIP(hipGetLastError());
  }

  template <typename T>
  void launch_sweep(const T *in, T *out, int nt, int d, const int n_in[3], int m_out,
                    const int *cnt, const int *col, const T *w, int K, bool add, hipStream_t s)
  {
    long o[3] = {n_in[0], n_in[1], n_in[2]};
    o[d]      = m_out;
    const long total = o[0] * o[1] * o[2] * nt;
    if (total == 0)
      return;
    k_sweep<T><<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s>>>(
      in, out, nt, d, n_in[0], n_in[1], n_in[2], m_out, cnt, col, w, K, add ? 1 : 0);
    STFEM_HIP(hipGetLastError());
  }

  template void launch_time_transfer<double>(long, const double *, int, int, const double *,
                                             double *, bool, hipStream_t);
  template void launch_time_transfer<float>(long, const float *, int, int, const float *,
                                            float *, bool, hipStream_t);
  template void launch_sweep<double>(const double *, double *, int, int, const int *, int,
                                     const int *, const int *, const double *, int, bool,
                                     hipStream_t);
  template void launch_sweep<float>(const float *, float *, int, int, const int *, int,
                                    const int *, const int *, const float *, int, bool,
                                    hipStream_t);
} // namespace stfem
