// Cell-patch Vanka smoother for gfx950
// (PreconditionVanka::vmult, tests/include/gmg.h:476-514 of nlsmrg/dealii-stfem,
// fused with the relaxation update of PreconditionRelaxation).
//
//   u += omega * sum_c (1/valence) .* scatter_c( B_c^{-1} gather_c(rhs - Au) )
//
// B_c = Alpha (x) K[I_c,I_c] + Beta (x) M[I_c,I_c], local index k + i*n_c
// (space fastest, time outer) as the reference builds it (gmg.h:460-468).
// Like the reference, the patch gathers and scatters every node of the cell,
// Dirichlet nodes included (their rows of B_c are decoupled).  With Au = null
// and omega = 1 on a zeroed u this is PreconditionVanka::vmult.
//
// Explicit form: one workgroup per cell, B_c^{-1} stored column-major so a
// column is read by consecutive lanes (coalesced); the gathered patch sits
// in LDS.  Factored form: B_c^{-1} = (I x S)(blockdiag_k W_k)(I x S^T) with
// S^T M_c S = I, S^T K_c S = diag(lambda), W_k = (lambda_k Alpha + Beta)^{-1};
// S is staged into LDS with coalesced loads when it fits.  The valence
// (number of cells sharing a node) of a structured-grid node is a product of
// per-direction multiplicities and is computed on the fly.
#include "kernels.h"

#include <type_traits>

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

    PatchMap make_map(const GridDesc &g)
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

    template <typename T>
    __device__ __forceinline__ void gather(const PatchMap &M, int cx, int cy, int cz, int nt,
                                           int nloc, const T *__restrict__ rhs,
                                           const T *__restrict__ ax, T *xs)
    {
      const int nb = nt * nloc;
      for (int l = threadIdx.x; l < nb; l += blockDim.x)
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

    template <typename T>
    __global__ void __launch_bounds__(256)
      k_vanka_explicit(const PatchMap M, int nt, int nloc, const T *__restrict__ binv,
                       const T *__restrict__ rhs, const T *__restrict__ ax, T *__restrict__ u,
                       T omega)
    {
      extern __shared__ unsigned char smem_raw[];
      T *const   xs   = reinterpret_cast<T *>(smem_raw);
      const long cell = blockIdx.x;
      const int  nb   = nt * nloc;
      int        cx, cy, cz;
      M.cell_xyz(cell, cx, cy, cz);
      gather(M, cx, cy, cz, nt, nloc, rhs, ax, xs);
      __syncthreads();
      const T *B = binv + (size_t)cell * nb * nb;
      for (int row = threadIdx.x; row < nb; row += blockDim.x)
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

    // Factored apply, one workgroup per cell.  LDS: x, y (nt*nloc each) and,
    // when STAGE, S (nloc*nloc).
    template <typename T, bool STAGE>
    __global__ void __launch_bounds__(256)
      k_vanka_factored(const PatchMap M, int nt, int nloc, const T *__restrict__ S,
                       const T *__restrict__ W, const T *__restrict__ rhs,
                       const T *__restrict__ ax, T *__restrict__ u, T omega)
    {
      extern __shared__ unsigned char smem_raw[];
      T *const   xs   = reinterpret_cast<T *>(smem_raw);
      T *const   ys   = xs + nt * nloc;
      T *const   ss   = ys + nt * nloc;
      const long cell = blockIdx.x;
      int        cx, cy, cz;
      M.cell_xyz(cell, cx, cy, cz);
      const T *Sg = S + (size_t)cell * nloc * nloc;
      if constexpr (STAGE)
        for (int l = threadIdx.x; l < nloc * nloc; l += blockDim.x)
          ss[l] = Sg[l];
      gather(M, cx, cy, cz, nt, nloc, rhs, ax, xs);
      __syncthreads();
      const T *Sc = STAGE ? ss : Sg;
      const T *Wc = W + (size_t)cell * nloc * nt * nt;
      // y_{i,k} = sum_a S[a][k] x_{i,a}  then  z_{:,k} = W_k y_{:,k}
      for (int k = threadIdx.x; k < nloc; k += blockDim.x)
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
      for (int a = threadIdx.x; a < nloc; a += blockDim.x)
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
    template <typename T, int NL, int NT>
    __global__ void __launch_bounds__(256)
      k_vanka_factored_v2(const PatchMap M, const T *__restrict__ S, const T *__restrict__ W,
                          const T *__restrict__ rhs, const T *__restrict__ ax,
                          T *__restrict__ u, T omega)
    {
      constexpr int G  = ((NL + 63) / 64) * 64;
      constexpr int P  = 256 / G;
      constexpr int LD = NL + 1;
      constexpr int CH = (NL + P - 1) / P; // contraction slice per group
      extern __shared__ unsigned char smem_raw[];
      T *const ss  = reinterpret_cast<T *>(smem_raw); // NL * LD
      T *const xs  = ss + NL * LD;                    // NT * NL
      T *const ys  = xs + NT * NL;                    // NT * NL
      T *const red = ys + NT * NL;                    // P * NT * NL

      const long cell = blockIdx.x;
      int        cx, cy, cz;
      M.cell_xyz(cell, cx, cy, cz);
      const T *Sg = S + (size_t)cell * NL * NL;
      for (int l = threadIdx.x; l < NL * NL; l += 256)
        {
          const int a = l / NL, k = l - a * NL;
          ss[a * LD + k] = Sg[l];
        }
      for (int l = threadIdx.x; l < NT * NL; l += 256)
        {
          const int  i = l / NL, j = l - i * NL;
          float      iv;
          const long g = M.node(cx, cy, cz, j, iv);
          T          v = rhs[g * NT + i];
          if (ax)
            v -= ax[g * NT + i];
          xs[l] = v;
        }
      __syncthreads();
      const int part = threadIdx.x / G, lane = threadIdx.x - part * G;
      const int a0 = part * CH, a1 = (a0 + CH < NL) ? a0 + CH : NL;
      // partial y_{i,k} over a in [a0, a1)
      if (lane < NL)
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
      for (int k = threadIdx.x; k < NL; k += 256)
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
      if (lane < NL)
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
      for (int l = threadIdx.x; l < NT * NL; l += 256)
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

    template <int NL>
    constexpr int v2_groups()
    {
      return 256 / (((NL + 63) / 64) * 64);
    }
    template <typename T, int NL, int NT>
    size_t v2_lds()
    {
      return sizeof(T) * (size_t(NL) * (NL + 1) + 2 * NT * NL + v2_groups<NL>() * NT * NL);
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
    const int nloc = (g.dim == 3) ? (g.p + 1) * (g.p + 1) * (g.p + 1) : (g.p + 1) * (g.p + 1);
    const int nb   = nt * nloc;
    const int thr  = nb >= 256 ? 256 : ((nb + 63) / 64) * 64;
    const size_t sm = size_t(nb) * sizeof(T);
    STFEM_REQUIRE(sm <= 64 * 1024, "vanka: patch too large for LDS");
    k_vanka_explicit<T><<<dim3((unsigned)g.n_cells), dim3((unsigned)thr), sm, s>>>(
      make_map(g), nt, nloc, binv, rhs, ax, u, omega);
    STFEM_HIP(hipGetLastError());
  }

  template <typename T>
  void launch_vanka_factored(const GridDesc &g, int nt, const T *S, const T *W, const T *rhs,
                             const T *ax, T *u, T omega, hipStream_t s)
  {
    if (g.n_cells == 0)
      return;
    STFEM_REQUIRE(nt <= 16, "vanka factored: at most 16 time dofs per patch");
    const int    nloc = (g.dim == 3) ? (g.p + 1) * (g.p + 1) * (g.p + 1) : (g.p + 1) * (g.p + 1);
    const int    thr  = nloc >= 256 ? 256 : ((nloc + 63) / 64) * 64;
    const size_t base = 2 * size_t(nt) * nloc * sizeof(T);
    const size_t full = base + size_t(nloc) * nloc * sizeof(T);
    const PatchMap M  = make_map(g);
    // compile-time patch shapes first
    bool done = false;
    auto try_v2 = [&](auto nl_c, auto nt_c) {
      constexpr int NL = decltype(nl_c)::value, NT = decltype(nt_c)::value;
      if (done || nloc != NL || nt != NT)
        return;
      const size_t lds = v2_lds<T, NL, NT>();
      if (lds > max_lds_per_block())
        return;
      if (lds > 64 * 1024)
        STFEM_HIP(hipFuncSetAttribute((const void *)k_vanka_factored_v2<T, NL, NT>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
      k_vanka_factored_v2<T, NL, NT><<<dim3((unsigned)g.n_cells), dim3(256), lds, s>>>(
        M, S, W, rhs, ax, u, omega);
      done = true;
    };
    auto for_nt = [&](auto nl_c) {
      try_v2(nl_c, std::integral_constant<int, 1>{});
      try_v2(nl_c, std::integral_constant<int, 2>{});
      try_v2(nl_c, std::integral_constant<int, 3>{});
      try_v2(nl_c, std::integral_constant<int, 4>{});
      try_v2(nl_c, std::integral_constant<int, 6>{});
      try_v2(nl_c, std::integral_constant<int, 8>{});
    };
    for_nt(std::integral_constant<int, 9>{});
    for_nt(std::integral_constant<int, 16>{});
    for_nt(std::integral_constant<int, 25>{});
    for_nt(std::integral_constant<int, 27>{});
    for_nt(std::integral_constant<int, 64>{});
    for_nt(std::integral_constant<int, 125>{});
    if (done)
      {
        STFEM_HIP(hipGetLastError());
        return;
      }
    if (full <= max_lds_per_block())
      {
        if (full > 64 * 1024)
          STFEM_HIP(hipFuncSetAttribute((const void *)k_vanka_factored<T, true>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, int(full)));
        k_vanka_factored<T, true><<<dim3((unsigned)g.n_cells), dim3((unsigned)thr), full, s>>>(
          M, nt, nloc, S, W, rhs, ax, u, omega);
      }
    else
      k_vanka_factored<T, false><<<dim3((unsigned)g.n_cells), dim3((unsigned)thr), base, s>>>(
        M, nt, nloc, S, W, rhs, ax, u, omega);
    STFEM_HIP(hipGetLastError());
  }

  template void launch_vanka_explicit<double>(const GridDesc &, int, const double *,
                                              const double *, const double *, double *, double,
                                              hipStream_t);
  template void launch_vanka_explicit<float>(const GridDesc &, int, const float *, const float *,
                                             const float *, float *, float, hipStream_t);
  template void launch_vanka_factored<double>(const GridDesc &, int, const double *,
                                              const double *, const double *, const double *,
                                              double *, double, hipStream_t);
  template void launch_vanka_factored<float>(const GridDesc &, int, const float *, const float *,
                                             const float *, const float *, float *, float,
                                             hipStream_t);
} // namespace stfem
