// Multigrid transfer kernels for gfx950, one launch per transfer step.
//
// Time transfers (MGTwoLevelTransferST, tests/include/gmg.h:121-135 of
// nlsmrg/dealii-stfem) are one streaming pass over the node-major vector;
// space transfers (deal.II MGTwoLevelTransfer built at gmg.h:371-380) are one
// 1D banded sweep per direction.  Per-entry device code is in dev_transfer.h.
#include "dev_transfer.h"

namespace stfem
{
  namespace
  {
    template <typename T>
    __global__ void k_time_transfer(long total, const T *__restrict__ M, int ntd, int nts,
                                    const T *__restrict__ src, T *__restrict__ dst, int add)
    {
      const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (idx < total)
        time_transfer_item(idx, M, ntd, nts, src, dst, add);
    }

    template <typename T>
    __global__ void k_sweep(long total, const T *__restrict__ in, T *__restrict__ out, int nt,
                            int d, int n0, int n1, int m, const int *__restrict__ cnt,
                            const int *__restrict__ col, const T *__restrict__ w, int K, int add)
    {
      const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (idx < total)
        sweep_item(idx, in, out, nt, d, n0, n1, m, cnt, col, w, K, add);
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
      total, M, ntd, nts, src, dst, add ? 1 : 0);
    STFEM_HIP(hipGetLastError());
  }

  template <typename T>
  void launch_sweep(const T *in, T *out, int nt, int d, const int n_in[3], int m_out,
                    const int *cnt, const int *col, const T *w, int K, bool add, hipStream_t s)
  {
    const long total = sweep_total(nt, d, n_in, m_out);
    if (total == 0)
      return;
    k_sweep<T><<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s>>>(
      total, in, out, nt, d, n_in[0], n_in[1], m_out, cnt, col, w, K, add ? 1 : 0);
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
