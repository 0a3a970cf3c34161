// Factored Vanka sweep (PreconditionVanka::vmult, tests/include/gmg.h:476-514
// of nlsmrg/dealii-stfem) for FP32 patches of 64 or 125 nodes with the next
// cell's eigenvector block S prefetched into LDS by gfx950's direct
// global-to-LDS loads.
//
// The sweep reads s (n_c^2 + n_c) bytes per cell once and spends about
// 4 flop per byte on them, so what bounds it is how much of S is in flight
// per CU.  The register-tiled kernel (vanka_v3_range) holds one cell's S in
// VGPRs and alternates load and product phases inside a workgroup.  Here
// two persistent workgroups per CU each walk a contiguous range of cells
// (x-neighbours, so that the rhs/ax gathers and the rows of the cell-major
// result buffer share cache lines): a workgroup copies cell c's S from its
// LDS stage into the same register tiles (thread (ta, tk) of a 16 x 16 grid
// holds S[ta + 16 m][4 (tk + 16 h) + e]), then immediately starts the
// global_load_lds_dwordx4 of cell c+1 into the stage (16 bytes per lane
// land at M0 + 16 * lane, no registers involved) and computes cell c while
// it streams.  The residual values and eigenvalues of the next cell are
// prefetched into registers the same way.
//
// The row partials of Sᵀ x are summed over the four row groups of a wave
// with the gfx950 lane swaps (v_permlane16/32_swap) and over the waves in
// LDS; the column partials of S z are summed within lane quads by DPP and
// over the quads in LDS, which keeps the LDS of a workgroup at 76 KB (two
// per CU).
#pragma once

#include "dev_vanka.h"

namespace stfem
{
  namespace
  {
    template <int NL, int NT>
    struct V5
    {
      using B                          = V3<float, NL, NT>;
      static constexpr int    VW       = 4;
      static constexpr int    LDG      = B::LDG;   // row stride of S in memory (floats)
      static constexpr int    R        = B::R;     // rows per thread
      static constexpr int    H        = B::H;     // column groups per thread
      static constexpr int    KP       = B::KP;    // padded column count
      static constexpr int    AP       = B::AP;    // padded row count
      static constexpr int    NX       = B::NX;    // residual values per thread
      static constexpr int    NCH      = LDG / 4;  // 16-byte chunks per row
      static constexpr int    CHUNKS   = NL * NCH; // chunks per cell
      static constexpr int    NI       = (CHUNKS + 255) / 256;
      static constexpr int    STAGE    = CHUNKS * 4;
      static constexpr int    RED      = 4 * NT * (KP > AP ? KP : AP);
      static constexpr size_t lds =
        sizeof(float) * (size_t(STAGE) + size_t(NT) * NL + size_t(NT) * KP + RED);
      static_assert(NCH == 16 * H, "a row is 16 x H chunks");
    };

    __device__ __forceinline__ float v5_pair_sum16(float x)
    {
      // x[l] + x[l ^ 16]: exchange odd and even rows of 16 lanes
      const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x),
                                                      false, false);
      return __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    __device__ __forceinline__ float v5_pair_sum32(float x)
    {
      // x[l] + x[l ^ 32]: exchange the halves of the wave
      const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x),
                                                      false, false);
      return __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    __device__ __forceinline__ float v5_quad_sum(float x)
    {
      // sum over the lane quad (xor 1, then xor 2)
      x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));
      x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x4E, 0xF, 0xF, false));
      return x;
    }

    template <int NL, int NT>
    __device__ __forceinline__ void v5_issue_s(const float *__restrict__ Sc, float *stage, int t)
    {
      using C = V5<NL, NT>;
      typedef __attribute__((address_space(3))) void lds_void;
      const int w = t >> 6;
#pragma unroll
      for (int j = 0; j < C::NI; ++j)
        {
          const int p = j * 256 + t; // chunk p of the cell, same position in LDS
          if (p < C::CHUNKS)
            __builtin_amdgcn_global_load_lds((const void *)(Sc + 4 * p),
                                             (lds_void *)(stage + 4 * (j * 256 + w * 64)), 16, 0,
                                             0);
        }
    }

    template <int NL, int NT>
    __device__ __forceinline__ void v5_prefetch_x(const PatchMap &M, long cell,
                                                  const float *__restrict__ rhs,
                                                  const float *__restrict__ ax,
                                                  const float *__restrict__ lam, int t,
                                                  float (&x)[V5<NL, NT>::NX], float &lk)
    {
      using C = V5<NL, NT>;
      int cx, cy, cz;
      M.cell_xyz(cell, cx, cy, cz);
#pragma unroll
      for (int j = 0; j < C::NX; ++j)
        {
          const int l = t + 256 * j;
          float     v = 0.0f;
          if (l < NT * NL)
            {
              const int  a = l / NT, i = l - a * NT;
              float      iv;
              const long gn = node_ct<NL>(M, cx, cy, cz, a, iv);
              v             = rhs[gn * NT + i];
              if (ax)
                v -= ax[gn * NT + i];
            }
          x[j] = v;
        }
      lk = (t < NL) ? lam[(size_t)cell * NL + t] : 0.0f;
    }

    // cells [c0, c1) of this workgroup; (omega / valence) B_c^{-1} (rhs - ax)|_c
    // into the cell-major buffer
    template <int NL, int NT>
    __device__ void vanka_v5_range(const PatchMap &M, long c0, long c1, const float *__restrict__ S,
                                   const float *__restrict__ lam, const float *__restrict__ talpha,
                                   const float *__restrict__ tbeta, const float *__restrict__ rhs,
                                   const float *__restrict__ ax, float *__restrict__ cells,
                                   float omega, int t, float *smem)
    {
      using C  = V5<NL, NT>;
      using Sh = V2Shape<NL>;
      constexpr int n = Sh::n, R = C::R, H = C::H, KP = C::KP, AP = C::AP, NX = C::NX,
                    LDG = C::LDG;
      typedef float vec_t __attribute__((ext_vector_type(4)));
      float *const stage = smem;
      float *const xs    = stage + C::STAGE; // [a][i]
      float *const ys    = xs + NT * NL;     // [i][k], row stride KP
      float *const red   = ys + NT * KP;     // 4 x NT x max(KP, AP)
      const int    ta = t >> 4, tk = t & 15, w = t >> 6;
      if (c0 >= c1)
        return;
      v5_issue_s<NL, NT>(S + (size_t)c0 * NL * LDG, stage, t);
      float xr[NX], lk;
      v5_prefetch_x<NL, NT>(M, c0, rhs, ax, lam, t, xr, lk);
      vec_t s[R][H];
      for (long c = c0; c < c1; ++c)
        {
          // cell c's S has landed (each wave waits for its own loads, the
          // barrier for everyone's)
          __builtin_amdgcn_s_waitcnt(0x0f70); // vmcnt(0)
          __syncthreads();
#pragma unroll
          for (int m = 0; m < R; ++m)
#pragma unroll
            for (int h = 0; h < H; ++h)
              {
                const int a = ta + 16 * m;
                s[m][h]     = a < NL ? *reinterpret_cast<const vec_t *>(
                                         stage + a * LDG + 4 * (tk + 16 * h)) :
                                       vec_t(0.0f);
              }
#pragma unroll
          for (int j = 0; j < NX; ++j)
            {
              const int l = t + 256 * j;
              if (l < NT * NL)
                xs[l] = xr[j];
            }
          const float lcur = lk;
          // everyone has read the stage: stream the next cell into it
          __syncthreads();
          if (c + 1 < c1)
            {
              v5_issue_s<NL, NT>(S + (size_t)(c + 1) * NL * LDG, stage, t);
              v5_prefetch_x<NL, NT>(M, c + 1, rhs, ax, lam, t, xr, lk);
            }

          // y = Sᵀ x over this thread's rows, summed over the wave's four
          // row groups (lanes l, l^16, l^32, l^48 share tk)
          {
            float y[H][4][NT];
#pragma unroll
            for (int h = 0; h < H; ++h)
#pragma unroll
              for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int i = 0; i < NT; ++i)
                  y[h][e][i] = 0.0f;
#pragma unroll
            for (int m = 0; m < R; ++m)
              {
                const int a  = ta + 16 * m;
                const int ac = a < NL ? a : NL - 1; // rows past NL hold zeros
                float     xv[NT];
#pragma unroll
                for (int i = 0; i < NT; ++i)
                  xv[i] = xs[ac * NT + i];
#pragma unroll
                for (int h = 0; h < H; ++h)
#pragma unroll
                  for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int i = 0; i < NT; ++i)
                      y[h][e][i] += s[m][h][e] * xv[i];
              }
#pragma unroll
            for (int h = 0; h < H; ++h)
#pragma unroll
              for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int i = 0; i < NT; ++i)
                  y[h][e][i] = v5_pair_sum32(v5_pair_sum16(y[h][e][i]));
            if ((t & 63) < 16)
#pragma unroll
              for (int h = 0; h < H; ++h)
#pragma unroll
                for (int i = 0; i < NT; ++i)
                  {
                    vec_t v;
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                      v[e] = y[h][e][i];
                    *reinterpret_cast<vec_t *>(red + (w * NT + i) * KP + 4 * (tk + 16 * h)) = v;
                  }
          }
          __syncthreads();
          // z_k = (lam_k Alpha + Beta)^{-1} y_k per eigenmode k
          if (t < KP)
            {
              float z[NT];
              if (t < NL)
                {
                  float y[NT];
#pragma unroll
                  for (int i = 0; i < NT; ++i)
                    y[i] = red[(0 * NT + i) * KP + t] + red[(1 * NT + i) * KP + t] +
                           red[(2 * NT + i) * KP + t] + red[(3 * NT + i) * KP + t];
                  solve_time_block<float, NT>(lcur, talpha, tbeta, y, z);
                }
              else
#pragma unroll
                for (int i = 0; i < NT; ++i)
                  z[i] = 0.0f; // padding columns
#pragma unroll
              for (int i = 0; i < NT; ++i)
                ys[i * KP + t] = z[i];
            }
          __syncthreads();
          // out = S z over this thread's columns, summed over lane quads
          {
            float o[R][NT];
#pragma unroll
            for (int m = 0; m < R; ++m)
#pragma unroll
              for (int i = 0; i < NT; ++i)
                o[m][i] = 0.0f;
#pragma unroll
            for (int h = 0; h < H; ++h)
              {
                vec_t zv[NT];
#pragma unroll
                for (int i = 0; i < NT; ++i)
                  zv[i] = *reinterpret_cast<const vec_t *>(ys + i * KP + 4 * (tk + 16 * h));
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                  for (int m = 0; m < R; ++m)
#pragma unroll
                    for (int i = 0; i < NT; ++i)
                      o[m][i] += s[m][h][e] * zv[i][e];
              }
#pragma unroll
            for (int m = 0; m < R; ++m)
#pragma unroll
              for (int i = 0; i < NT; ++i)
                o[m][i] = v5_quad_sum(o[m][i]);
            if ((tk & 3) == 0)
#pragma unroll
              for (int m = 0; m < R; ++m)
#pragma unroll
                for (int i = 0; i < NT; ++i)
                  red[((tk >> 2) * NT + i) * AP + ta + 16 * m] = o[m][i];
          }
          __syncthreads();
          // sum the quads' partials, scale by omega / valence, store
#pragma unroll
          for (int j = 0; j < (NT * AP + 255) / 256; ++j)
            {
              const int l = t + 256 * j;
              const int a = l % AP, i = l / AP;
              if (i < NT && a < NL)
                {
                  const float acc = red[(0 * NT + i) * AP + a] + red[(1 * NT + i) * AP + a] +
                                    red[(2 * NT + i) * AP + a] + red[(3 * NT + i) * AP + a];
                  int cx, cy, cz;
                  M.cell_xyz(c, cx, cy, cz);
                  float iv;
                  (void)node_ct<NL>(M, cx, cy, cz, a, iv);
                  const int la = a % n, lb = (a / n) % n, lc = (Sh::dim == 3) ? a / (n * n) : 0;
                  cells[cell_slot<Sh::dim, n>(M.nc0, M.nc1, NT, cx, cy, cz, la, lb, lc, i)] =
                    omega * iv * acc;
                }
            }
        }
    }

    template <int NL, int NT>
    __global__ void __launch_bounds__(256, 2)
      k_vanka_factored_v5(const PatchMap M, long n_cells, const float *__restrict__ S,
                          const float *__restrict__ lam, const float *__restrict__ ta,
                          const float *__restrict__ tb, const float *__restrict__ rhs,
                          const float *__restrict__ ax, float *__restrict__ cells, float omega)
    {
      extern __shared__ __align__(16) unsigned char smem_raw[];
      const long per = (n_cells + gridDim.x - 1) / gridDim.x;
      const long c0  = (long)blockIdx.x * per;
      const long c1  = (c0 + per < n_cells) ? c0 + per : n_cells;
      vanka_v5_range<NL, NT>(M, c0, c1, S, lam, ta, tb, rhs, ax, cells, omega, (int)threadIdx.x,
                             reinterpret_cast<float *>(smem_raw));
    }
  } // namespace
} // namespace stfem
