// Cell-patch Vanka smoother for gfx950, one launch per sweep
// (PreconditionVanka::vmult, tests/include/gmg.h:476-514 of nlsmrg/dealii-stfem,
// fused with the relaxation update of PreconditionRelaxation).
//
//   u += omega * sum_c (1/valence) .* scatter_c( B_c^{-1} gather_c(rhs - Au) )
//
// With Au = null and omega = 1 on a zeroed u this is PreconditionVanka::vmult.
// Explicit form: one workgroup per cell, B_c^{-1} stored column-major so a
// column is read by consecutive lanes (coalesced); the gathered patch sits in
// LDS.  Factored form: B_c^{-1} = (I x S)(blockdiag_k W_k)(I x S^T) with
// S^T M_c S = I, S^T K_c S = diag(lambda), W_k = (lambda_k Alpha + Beta)^{-1}
// applied by solving the nt x nt time block from lambda_k.  The per-cell device code is in
// dev_vanka.h.
#include <mutex>
#include <vector>

#include "dev_vanka.h"
#include "dev_vanka_dma.h"

#include <algorithm>
#include <type_traits>

namespace stfem
{
  namespace
  {
    template <typename T>
    __global__ void __launch_bounds__(256)
      k_vanka_explicit(const PatchMap M, int nt, int nloc, const T *__restrict__ binv,
                       const T *__restrict__ rhs, const T *__restrict__ ax, T *__restrict__ u,
                       T omega)
    {
      extern __shared__ unsigned char smem_raw[];
      vanka_explicit_cell<T>(M, nt, nloc, binv, rhs, ax, u, omega,
                             xcd_block((long)blockIdx.x, (long)gridDim.x), true,
                             (int)threadIdx.x, (int)blockDim.x,
                             reinterpret_cast<T *>(smem_raw));
    }

    template <typename T, bool STAGE>
    __global__ void __launch_bounds__(256)
      k_vanka_factored(const PatchMap M, int nt, int nloc, int ld, const T *__restrict__ S,
                       const T *__restrict__ lam, const T *__restrict__ ta,
                       const T *__restrict__ tb, const T *__restrict__ rhs,
                       const T *__restrict__ ax, T *__restrict__ u, T omega)
    {
      extern __shared__ unsigned char smem_raw[];
      vanka_factored_cell<T, STAGE>(M, nt, nloc, ld, S, lam, ta, tb, rhs, ax, u, omega,
                                    xcd_block((long)blockIdx.x, (long)gridDim.x),
                                    (int)threadIdx.x, (int)blockDim.x,
                                    reinterpret_cast<T *>(smem_raw));
    }

    template <typename T, int NL, int NT, bool TWO_PASS>
    __global__ void __launch_bounds__(256)
      k_vanka_factored_v2(const PatchMap M, const T *__restrict__ S, const T *__restrict__ lam,
                          const T *__restrict__ ta, const T *__restrict__ tb,
                          const T *__restrict__ rhs, const T *__restrict__ ax,
                          T *__restrict__ u, T *__restrict__ cells, T omega)
    {
      extern __shared__ __align__(16) unsigned char smem_raw[];
      vanka_v2_cell<T, NL, NT, TWO_PASS>(M, S, lam, ta, tb, rhs, ax, u, cells, omega,
                                         xcd_block((long)blockIdx.x, (long)gridDim.x),
                                         (int)threadIdx.x, reinterpret_cast<T *>(smem_raw));
    }

    // !PIPE: register budget for four (FP32) or two (FP64) workgroups per CU
    template <typename T, int NL, int NT, bool PIPE, bool NTS = true>
    __global__ void __launch_bounds__(256, PIPE ? 1 : (sizeof(T) == 4 ? 4 : 2))
      k_vanka_factored_v3(const PatchMap M, long n_cells, const T *__restrict__ S,
                          const T *__restrict__ lam, const T *__restrict__ ta,
                          const T *__restrict__ tb, const T *__restrict__ rhs,
                          const T *__restrict__ ax, T *__restrict__ cells, T omega)
    {
      extern __shared__ __align__(16) unsigned char smem_raw[];
      vanka_v3_range<T, NL, NT, PIPE, false, NTS>(M, n_cells, S, lam, ta, tb, rhs, ax, cells,
                                                  omega, (int)threadIdx.x,
                                                  reinterpret_cast<T *>(smem_raw));
    }
    // cell buffers above this size take non-temporal stores of the sweep's
    // results (up to it they stay in the 256 MB last-level cache for the
    // node gather that follows)
    constexpr size_t NT_STORE_BYTES = size_t(256) << 20;
    // the same sweep with A x summed from the operator's cell buffer opc
    template <typename T, int NL, int NT>
    __global__ void __launch_bounds__(256, sizeof(T) == 4 ? 4 : 2)
      k_vanka_factored_v3_opc(const PatchMap M, const GridDesc g, const T *__restrict__ S,
                              const T *__restrict__ lam, const T *__restrict__ ta,
                              const T *__restrict__ tb, const T *__restrict__ rhs,
                              const T *__restrict__ opc, T *__restrict__ cells, T omega)
    {
      extern __shared__ __align__(16) unsigned char smem_raw[];
      vanka_v3_range<T, NL, NT, false, true>(M, g.n_cells, S, lam, ta, tb, rhs, nullptr, cells,
                                             omega, (int)threadIdx.x,
                                             reinterpret_cast<T *>(smem_raw), &g, opc);
    }

    size_t max_lds_per_block()
    {
      static size_t v = 0;
      if (!v)
        {
          int dev = 0, bytes = 0;
          STFEM_HIP(hipGetDevice(&dev));
          STFEM_HIP(hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
          v = size_t(bytes);
        }
      return v;
    }
  } // namespace

  template <typename T>
  void launch_vanka_explicit(const GridDesc &g, int nt, const T *binv, const T *rhs, const T *ax,
                             T *u, T omega, hipStream_t s)
  {
    if (g.n_cells == 0)
      return;
    const int nloc = patch_size(g);
    const int nb   = nt * nloc;
    const int thr  = nb >= 256 ? 256 : ((nb + 63) / 64) * 64;
    const size_t sm = size_t(nb) * sizeof(T);
    STFEM_REQUIRE(sm <= 64 * 1024, "vanka: patch too large for LDS");
    k_vanka_explicit<T><<<dim3((unsigned)g.n_cells), dim3((unsigned)thr), sm, s>>>(
      make_map(g), nt, nloc, binv, rhs, ax, u, omega);
    STFEM_HIP(hipGetLastError());
  }

  namespace
  {
    // the compile-time patch shapes of the factored kernels, f(NL, NT)
    template <typename F>
    void for_each_shape(F &&f)
    {
      auto for_nt = [&](auto nl_c) {
        f(nl_c, std::integral_constant<int, 1>{});
        f(nl_c, std::integral_constant<int, 2>{});
        f(nl_c, std::integral_constant<int, 3>{});
        f(nl_c, std::integral_constant<int, 4>{});
        f(nl_c, std::integral_constant<int, 6>{});
        f(nl_c, std::integral_constant<int, 8>{});
      };
      for_nt(std::integral_constant<int, 4>{});
      for_nt(std::integral_constant<int, 8>{});
      for_nt(std::integral_constant<int, 9>{});
      for_nt(std::integral_constant<int, 16>{});
      for_nt(std::integral_constant<int, 25>{});
      for_nt(std::integral_constant<int, 27>{});
      for_nt(std::integral_constant<int, 64>{});
      for_nt(std::integral_constant<int, 125>{});
      for_nt(std::integral_constant<int, 36>{});
    }
  } // namespace

  template <typename T>
  bool vanka_fused_residual_supported(const GridDesc &g, int nt)
  {
    const int nloc = patch_size(g);
    bool      ok   = false;
    for_each_shape([&](auto nl_c, auto nt_c) {
      constexpr int NL = decltype(nl_c)::value, NT = decltype(nt_c)::value;
      if constexpr (NL <= 128 && V3<T, NL, NT>::lds <= 80 * 1024)
        if (nloc == NL && nt == NT && V3<T, NL, NT>::lds <= max_lds_per_block())
          ok = true;
    });
    return ok;
  }

  namespace
  {
    // the FP32 sweep with LDS-prefetched S (dev_vanka_dma.h): two persistent
    // workgroups per CU; false if the shape has no such kernel
    bool launch_vanka_v5(const GridDesc &g, int nt, const float *S, const float *lam,
                         const float *ta, const float *tb, const float *rhs, const float *ax,
                         float *cells, float omega, hipStream_t s)
    {
      const int nloc = patch_size(g);
      if (g.dim != 3 || nt < 1 || nt > 4 || (nloc != 64 && nloc != 125))
        return false;
      static std::mutex       mtx;
      static std::vector<int> cus;
      int                     dev = 0, ncu = 0;
      STFEM_HIP(hipGetDevice(&dev));
      {
        std::lock_guard<std::mutex> lk(mtx);
        if (dev >= int(cus.size()))
          cus.resize(dev + 1, 0);
        if (!cus[dev])
          STFEM_HIP(hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev));
        ncu = cus[dev];
      }
      const PatchMap M    = make_map(g);
      const long     grid = std::min<long>(g.n_cells, 2L * std::max(1, ncu));
      bool           done = false;
      auto           go   = [&](auto nl_c, auto nt_c) {
        constexpr int NL = decltype(nl_c)::value, NT = decltype(nt_c)::value;
        if (done || nloc != NL || nt != NT)
          return;
        constexpr size_t lds = V5<NL, NT>::lds;
        if (lds > max_lds_per_block())
          return;
        if (lds > 64 * 1024)
          STFEM_HIP(hipFuncSetAttribute((const void *)k_vanka_factored_v5<NL, NT>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        k_vanka_factored_v5<NL, NT><<<dim3((unsigned)grid), dim3(256), lds, s>>>(
          M, g.n_cells, S, lam, ta, tb, rhs, ax, cells, omega);
        STFEM_HIP(hipGetLastError());
        done = true;
      };
      auto for_nt = [&](auto nl_c) {
        go(nl_c, std::integral_constant<int, 1>{});
        go(nl_c, std::integral_constant<int, 2>{});
        go(nl_c, std::integral_constant<int, 3>{});
        go(nl_c, std::integral_constant<int, 4>{});
      };
      for_nt(std::integral_constant<int, 64>{});
      for_nt(std::integral_constant<int, 125>{});
      return done;
    }
  } // namespace

  template <typename T>
  void launch_vanka_factored(const GridDesc &g, int nt, const T *S, const T *lam, const T *ta,
                             const T *tb, const T *rhs, const T *ax, T *u, T omega, T *cells,
                             int mode, int variant, hipStream_t s, const T *opc)
  {
    if (g.n_cells == 0)
      return;
    STFEM_REQUIRE(!opc || cells, "vanka factored: the fused residual needs the two-pass form");
    STFEM_REQUIRE(mode == 0 || mode == 1 || (mode == 3 && cells),
                  "vanka factored: mode must be 0, 1 or (two-pass form) 3");
    STFEM_REQUIRE(nt <= 16, "vanka factored: at most 16 time dofs per patch");
    const int    nloc = patch_size(g);
    const int    thr  = nloc >= 256 ? 256 : ((nloc + 63) / 64) * 64;
    const size_t base = 2 * size_t(nt) * nloc * sizeof(T);
    const size_t full = base + size_t(nloc) * nloc * sizeof(T);
    const PatchMap M  = make_map(g);
    if constexpr (std::is_same_v<T, float>)
      if (variant == 5 && cells && !opc &&
          launch_vanka_v5(g, nt, S, lam, ta, tb, rhs, ax, cells, omega, s))
        {
          if (mode != 3)
            launch_gather_cells<T>(g, nt, cells, u, nullptr, mode, false, s);
          return;
        }
    // compile-time patch shapes first
    bool done = false;
    auto try_v3 = [&](auto nl_c, auto nt_c) {
      constexpr int NL = decltype(nl_c)::value, NT = decltype(nt_c)::value;
      if constexpr (NL <= 128 && V3<T, NL, NT>::lds <= 80 * 1024)
        {
          if (done || nloc != NL || nt != NT || !cells)
            return;
          const bool use = opc || variant == 3 || (variant == 4 && sizeof(T) == 4) ||
                           (variant == 0 && NL >= 64);
          if (!use)
            return;
          constexpr size_t lds = V3<T, NL, NT>::lds;
          if (lds > max_lds_per_block())
            return;
          if (lds > 64 * 1024)
            {
              STFEM_HIP(hipFuncSetAttribute((const void *)k_vanka_factored_v3<T, NL, NT, false>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
              STFEM_HIP(
                hipFuncSetAttribute((const void *)k_vanka_factored_v3<T, NL, NT, false, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
              STFEM_HIP(hipFuncSetAttribute((const void *)k_vanka_factored_v3<T, NL, NT, true>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
              STFEM_HIP(hipFuncSetAttribute((const void *)k_vanka_factored_v3_opc<T, NL, NT>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
            }
          if (opc)
            k_vanka_factored_v3_opc<T, NL, NT><<<dim3((unsigned)g.n_cells), dim3(256), lds, s>>>(
              M, g, S, lam, ta, tb, rhs, opc, cells, omega);
          else if (variant == 4)
            {
              // resident workgroups of the persistent form, per device
              static std::mutex       mtx;
              static std::vector<int> per_device;
              int                     dev = 0, resident = 0;
              STFEM_HIP(hipGetDevice(&dev));
              {
                std::lock_guard<std::mutex> lk(mtx);
                if (dev >= int(per_device.size()))
                  per_device.resize(dev + 1, 0);
                if (!per_device[dev])
                  {
                    int cus = 0, per = 0;
                    STFEM_HIP(
                      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
                    STFEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                      &per, (const void *)k_vanka_factored_v3<T, NL, NT, true>, 256, lds));
                    per_device[dev] = std::max(1, cus * std::max(1, per));
                  }
                resident = per_device[dev];
              }
              const long grid = std::min<long>(g.n_cells, resident);
              k_vanka_factored_v3<T, NL, NT, true><<<dim3((unsigned)grid), dim3(256), lds, s>>>(
                M, g.n_cells, S, lam, ta, tb, rhs, ax, cells, omega);
            }
          else if (size_t(g.n_cells) * V3<T, NL, NT>::CS * sizeof(T) > NT_STORE_BYTES)
            k_vanka_factored_v3<T, NL, NT, false><<<dim3((unsigned)g.n_cells), dim3(256), lds, s>>>(
              M, g.n_cells, S, lam, ta, tb, rhs, ax, cells, omega);
          else
            k_vanka_factored_v3<T, NL, NT, false, false>
              <<<dim3((unsigned)g.n_cells), dim3(256), lds, s>>>(M, g.n_cells, S, lam, ta, tb, rhs,
                                                                 ax, cells, omega);
          STFEM_HIP(hipGetLastError());
          if (mode != 3)
            launch_gather_cells<T>(g, NT, cells, u, nullptr, mode, false, s, V3<T, NL, NT>::CS);
          done = true;
        }
    };
    auto try_v2 = [&](auto nl_c, auto nt_c) {
      constexpr int NL = decltype(nl_c)::value, NT = decltype(nt_c)::value;
      try_v3(nl_c, nt_c);
      if (done || nloc != NL || nt != NT)
        return;
      const size_t lds = v2_lds<T, NL, NT>();
      if (lds > max_lds_per_block())
        return;
      if (cells)
        {
          if (lds > 64 * 1024)
            STFEM_HIP(hipFuncSetAttribute((const void *)k_vanka_factored_v2<T, NL, NT, true>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
          k_vanka_factored_v2<T, NL, NT, true><<<dim3((unsigned)g.n_cells), dim3(256), lds, s>>>(
            M, S, lam, ta, tb, rhs, ax, u, cells, omega);
          STFEM_HIP(hipGetLastError());
          if (mode != 3)
            launch_gather_cells<T>(g, NT, cells, u, nullptr, mode, false, s);
        }
      else
        {
          if (mode == 0)
            launch_fill_zero<T>(u, size_t(g.n_nodes) * NT, s);
          if (lds > 64 * 1024)
            STFEM_HIP(hipFuncSetAttribute((const void *)k_vanka_factored_v2<T, NL, NT, false>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
          k_vanka_factored_v2<T, NL, NT, false><<<dim3((unsigned)g.n_cells), dim3(256), lds, s>>>(
            M, S, lam, ta, tb, rhs, ax, u, cells, omega);
        }
      done = true;
    };
    if (opc)
      {
        for_each_shape(try_v3);
        STFEM_REQUIRE(done, "vanka factored: no fused-residual kernel for this patch shape");
        STFEM_HIP(hipGetLastError());
        return;
      }
    for_each_shape(try_v2);
    if (done)
      {
        STFEM_HIP(hipGetLastError());
        return;
      }
    STFEM_REQUIRE(mode != 3, "vanka factored: no two-pass kernel for this patch shape");
    if (mode == 0)
      launch_fill_zero<T>(u, size_t(g.n_nodes) * nt, s);
    if (full <= max_lds_per_block())
      {
        if (full > 64 * 1024)
          STFEM_HIP(hipFuncSetAttribute((const void *)k_vanka_factored<T, true>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, int(full)));
        k_vanka_factored<T, true><<<dim3((unsigned)g.n_cells), dim3((unsigned)thr), full, s>>>(
          M, nt, nloc, vanka_s_ld(nloc, sizeof(T)), S, lam, ta, tb, rhs, ax, u, omega);
      }
    else
      k_vanka_factored<T, false><<<dim3((unsigned)g.n_cells), dim3((unsigned)thr), base, s>>>(
        M, nt, nloc, vanka_s_ld(nloc, sizeof(T)), S, lam, ta, tb, rhs, ax, u, omega);
    STFEM_HIP(hipGetLastError());
  }

  template void launch_vanka_explicit<double>(const GridDesc &, int, const double *,
                                              const double *, const double *, double *, double,
                                              hipStream_t);
  template void launch_vanka_explicit<float>(const GridDesc &, int, const float *, const float *,
                                             const float *, float *, float, hipStream_t);
  template void launch_vanka_factored<double>(const GridDesc &, int, const double *,
                                              const double *, const double *, const double *,
                                              const double *, const double *, double *, double,
                                              double *, int, int, hipStream_t, const double *);
  template void launch_vanka_factored<float>(const GridDesc &, int, const float *, const float *,
                                             const float *, const float *, const float *,
                                             const float *, float *, float, float *, int, int,
                                             hipStream_t, const float *);
  template bool vanka_fused_residual_supported<double>(const GridDesc &, int);
  template bool vanka_fused_residual_supported<float>(const GridDesc &, int);
} // namespace stfem
