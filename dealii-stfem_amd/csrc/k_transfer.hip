// Multigrid transfer kernels for gfx950, one launch per transfer step.
//
// Time transfers (MGTwoLevelTransferST, tests/include/gmg.h:121-135 of
// nlsmrg/dealii-stfem) are one streaming pass over the node-major vector;
// space transfers (deal.II MGTwoLevelTransfer built at gmg.h:371-380) are one
// 1D banded sweep per direction, one thread per node (all its time
// components) or per entry.  The device code is in dev_transfer.h.
#include "dev_transfer.h"

#include <cstdint>
#include <type_traits>

namespace stfem
{
  namespace
  {
    template <typename T, typename I>
    __global__ void k_time_transfer(I total, const T *__restrict__ M, int ntd, int nts,
                                    const T *__restrict__ src, T *__restrict__ dst, int add)
    {
      const I idx = I(blockIdx.x) * I(blockDim.x) + I(threadIdx.x);
      if (idx < total)
        time_transfer_item<T, I>(idx, M, ntd, nts, src, dst, add);
    }

    template <typename T, typename I, int KM>
    __global__ void k_sweep(I total, const T *__restrict__ in, T *__restrict__ out, int nt,
                            int d, int n0, int n1, int m, const int *__restrict__ cnt,
                            const int *__restrict__ col, const T *__restrict__ w, int K, int add)
    {
      const I idx = I(blockIdx.x) * I(blockDim.x) + I(threadIdx.x);
      if (idx < total)
        sweep_item<T, I, KM>(idx, in, out, nt, d, n0, n1, m, cnt, col, w, K, add);
    }

    template <typename T, typename I, int KM, int NT>
    __global__ void k_sweep_node(I nodes, const T *__restrict__ in, T *__restrict__ out, int d,
                                 int n0, int n1, int m, const int *__restrict__ cnt,
                                 const int *__restrict__ col, const T *__restrict__ w, int K,
                                 int add)
    {
      const I pt = I(blockIdx.x) * I(blockDim.x) + I(threadIdx.x);
      if (pt < nodes)
        sweep_node<T, I, KM, NT>(pt, in, out, d, n0, n1, m, cnt, col, w, K, add);
    }

    // sweeps along x (d = 0): the workgroup stages RPW whole input rows in
    // LDS (coalesced node-vector loads), then each thread computes output
    // nodes of those rows from LDS; taps summed in the order of sweep_node
    template <typename T, int KM, int NT, int RPW>
    __global__ void __launch_bounds__(256)
      k_sweep_x(long rows, const T *__restrict__ in, T *__restrict__ out, int n0, int m,
                const int *__restrict__ cnt, const int *__restrict__ col,
                const T *__restrict__ w, int K, int add)
    {
      extern __shared__ __align__(16) unsigned char smem_raw[];
      T *const   buf = reinterpret_cast<T *>(smem_raw); // RPW x n0 nodes x NT
      const long r0  = (long)blockIdx.x * RPW;
      const int  tid = threadIdx.x;
      for (int e = tid; e < RPW * n0; e += 256)
        {
          const int  rr = e / n0, i = e - rr * n0;
          const long r  = r0 + rr;
          if (r < rows)
            {
              T v[NT];
              load_comps<T, NT>(in + (r * n0 + i) * NT, v);
              store_comps<T, NT>(buf + e * NT, v);
            }
        }
      __syncthreads();
      for (int e = tid; e < RPW * m; e += 256)
        {
          const int  rr = e / m, o = e - rr * m;
          const long r  = r0 + rr;
          if (r >= rows)
            continue;
          const T *const row = buf + rr * n0 * NT;
          const int      c   = cnt[o];
          T              s[NT];
#pragma unroll
          for (int j = 0; j < NT; ++j)
            s[j] = T(0);
          if constexpr (KM == 0)
            {
#pragma unroll 4
              for (int k = 0; k < c; ++k)
                {
                  const T wk = w[o * K + k];
                  T       x[NT];
                  load_comps<T, NT>(row + col[o * K + k] * NT, x);
#pragma unroll
                  for (int j = 0; j < NT; ++j)
                    s[j] += wk * x[j];
                }
            }
          else
            {
              int cl[KM];
              T   wl[KM];
#pragma unroll
              for (int k = 0; k < KM; ++k)
                if (k < K)
                  {
                    cl[k] = col[o * K + k];
                    wl[k] = w[o * K + k];
                  }
#pragma unroll
              for (int k = 0; k < KM; ++k)
                if (k < c)
                  {
                    T x[NT];
                    load_comps<T, NT>(row + cl[k] * NT, x);
#pragma unroll
                    for (int j = 0; j < NT; ++j)
                      s[j] += wl[k] * x[j];
                  }
            }
          T *const dst = out + (r * m + o) * NT;
          if (add)
            {
              T v[NT];
              load_comps<T, NT>(dst, v);
#pragma unroll
              for (int j = 0; j < NT; ++j)
                s[j] = v[j] + s[j];
            }
          store_comps<T, NT>(dst, s);
        }
    }
    constexpr int    SWEEP_X_RPW = 2;          // input rows per workgroup
    constexpr size_t SWEEP_X_LDS = 48 * 1024;  // largest staged rows

    // one thread per node and all its nt (1-4) components where the vectors
    // allow the node's vector accesses; KM as for k_sweep
    template <typename T, typename I>
    bool launch_sweep_node(hipStream_t s, I nodes, const T *in, T *out, int nt, int d, int n0,
                           int n1, int m, const int *cnt, const int *col, const T *w, int K,
                           int add)
    {
      if (nt < 1 || nt > 4 || reinterpret_cast<uintptr_t>(in) % 16 != 0 ||
          reinterpret_cast<uintptr_t>(out) % 16 != 0)
        return false;
      const dim3   grid((unsigned)((nodes + 255) / 256));
      const size_t xlds = size_t(SWEEP_X_RPW) * n0 * nt * sizeof(T);
      auto go = [&](auto nt_c, auto km_c) {
        constexpr int NT = decltype(nt_c)::value, KM = decltype(km_c)::value;
        if (d == 0 && xlds <= SWEEP_X_LDS)
          {
            // rows of the output: nodes / m
            const long rows = long(nodes) / m;
            k_sweep_x<T, KM, NT, SWEEP_X_RPW>
              <<<dim3((unsigned)((rows + SWEEP_X_RPW - 1) / SWEEP_X_RPW)), dim3(256), xlds, s>>>(
                rows, in, out, n0, m, cnt, col, w, K, add);
          }
        else
          k_sweep_node<T, I, KM, NT><<<grid, dim3(256), 0, s>>>(nodes, in, out, d, n0, n1, m, cnt,
                                                                col, w, K, add);
      };
      auto by_k = [&](auto nt_c) {
        if (K <= 4)
          go(nt_c, std::integral_constant<int, 4>{});
        else if (K <= 8)
          go(nt_c, std::integral_constant<int, 8>{});
        else
          go(nt_c, std::integral_constant<int, 0>{});
      };
      switch (nt)
        {
          case 1: by_k(std::integral_constant<int, 1>{}); break;
          case 2: by_k(std::integral_constant<int, 2>{}); break;
          case 3: by_k(std::integral_constant<int, 3>{}); break;
          default: by_k(std::integral_constant<int, 4>{}); break;
        }
      return true;
    }

    // k_sweep with the smallest tap bound KM >= K for rows of up to 8 taps
    // (prolongation: p + 1); the wider restriction rows (up to 4p + 1 taps,
    // most rows using about half) run faster with the tap loop (KM = 0)
    template <typename T, typename I>
    void launch_sweep_k(dim3 grid, hipStream_t s, I total, const T *in, T *out, int nt, int d,
                        int n0, int n1, int m, const int *cnt, const int *col, const T *w, int K,
                        int add)
    {
      if (K <= 4)
        k_sweep<T, I, 4><<<grid, dim3(256), 0, s>>>(total, in, out, nt, d, n0, n1, m, cnt, col, w,
                                                     K, add);
      else if (K <= 8)
        k_sweep<T, I, 8><<<grid, dim3(256), 0, s>>>(total, in, out, nt, d, n0, n1, m, cnt, col, w,
                                                     K, add);
      else
        k_sweep<T, I, 0><<<grid, dim3(256), 0, s>>>(total, in, out, nt, d, n0, n1, m, cnt, col, w,
                                                     K, add);
    }

    // 32-bit indices are used when every index, and every thread index of the
    // rounded-up grid, stays below this
    constexpr long k_max_u32 = 4294967295L - 256;
    __global__ void k_wave_velocity(long n_nodes, int nts, int ni, int dg,
                                    const double *__restrict__ AixB,
                                    const double *__restrict__ AixG,
                                    const double *__restrict__ AixZ, const double *__restrict__ u,
                                    const double *__restrict__ prev_u,
                                    const double *__restrict__ prev_v, double *__restrict__ v)
    {
      const long node = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (node >= n_nodes)
        return;
      const int     nt = nts * ni;
      const double *un = u + node * nt;
      double       *vn = v + node * nt;
      for (int it = 0; it < nts; ++it)
        {
          const double pu = it == 0 ? prev_u[node] : un[it * ni - 1];
          const double pv = it == 0 ? prev_v[node] : vn[it * ni - 1];
          for (int i = 0; i < ni; ++i)
            {
              double s = 0.0;
              for (int j = 0; j < ni; ++j)
                s += AixB[i * ni + j] * un[it * ni + j];
              s += dg ? AixG[i] * pu : AixG[i] * pv + AixZ[i] * pu;
              vn[it * ni + i] = s;
            }
        }
    }
  } // namespace

  void launch_wave_velocity(long n_nodes, int nts, int ni, bool dg, const double *AixB,
                            const double *AixG, const double *AixZ, const double *u,
                            const double *prev_u, const double *prev_v, double *v, hipStream_t s)
  {
    if (n_nodes == 0)
      return;
    k_wave_velocity<<<dim3((unsigned)((n_nodes + 255) / 256)), 256, 0, s>>>(
      n_nodes, nts, ni, dg ? 1 : 0, AixB, AixG, AixZ, u, prev_u, prev_v, v);
    STFEM_HIP(hipGetLastError());
  }

  template <typename T>
  void launch_time_transfer(long n_nodes, const T *M, int ntd, int nts, const T *src, T *dst,
                            bool add, hipStream_t s)
  {
    const long total = n_nodes * ntd;
    if (total == 0)
      return;
    const dim3 grid((unsigned)((total + 255) / 256));
    if (n_nodes * std::max(ntd, nts) <= k_max_u32)
      k_time_transfer<T, unsigned><<<grid, dim3(256), 0, s>>>(unsigned(total), M, ntd, nts, src,
                                                             dst, add ? 1 : 0);
    else
      k_time_transfer<T, unsigned long><<<grid, dim3(256), 0, s>>>(total, M, ntd, nts, src, dst,
                                                                  add ? 1 : 0);
    STFEM_HIP(hipGetLastError());
  }

  template <typename T>
  void launch_sweep(const T *in, T *out, int nt, int d, const int n_in[3], int m_out,
                    const int *cnt, const int *col, const T *w, int K, bool add, hipStream_t s)
  {
    const long total = sweep_total(nt, d, n_in, m_out);
    if (total == 0)
      return;
    const long total_in = (long)n_in[0] * n_in[1] * n_in[2] * nt;
    const dim3 grid((unsigned)((total + 255) / 256));
    const bool u32 = std::max(total, total_in) <= k_max_u32;
    // one thread per node where possible, else one per entry
    if (u32)
      {
        if (!launch_sweep_node<T, unsigned>(s, unsigned(total / nt), in, out, nt, d, n_in[0],
                                            n_in[1], m_out, cnt, col, w, K, add ? 1 : 0))
          launch_sweep_k<T, unsigned>(grid, s, unsigned(total), in, out, nt, d, n_in[0], n_in[1],
                                      m_out, cnt, col, w, K, add ? 1 : 0);
      }
    else if (!launch_sweep_node<T, unsigned long>(s, total / nt, in, out, nt, d, n_in[0], n_in[1],
                                                  m_out, cnt, col, w, K, add ? 1 : 0))
      launch_sweep_k<T, unsigned long>(grid, s, total, in, out, nt, d, n_in[0], n_in[1], m_out,
                                       cnt, col, w, K, add ? 1 : 0);
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
