// Batched generalized symmetric eigensolver for the cell patches of the
// factored Vanka (gfx950, FP64): K_c z = lambda M_c z with z^T M_c z = 1,
// for patches of n <= 128 nodes.  PreconditionVanka (tests/include/gmg.h:
// 409-474 of nlsmrg/dealii-stfem) inverts the patch blocks
// B_c = Alpha (x) K_c + Beta (x) M_c; these kernels produce the eigenpairs
// that factor B_c^{-1} (the factored form of the smoother).
//
// Four launches per chunk of patches, each patch matrix held in LDS (one
// patch per CU; 1024 threads per workgroup, so that the matrix-vector
// products and rank-one updates of the sequential steps are spread over
// sixteen wavefronts):
//   1. k_eig_cholesky       one workgroup per patch: M = L L^T (right-looking,
//                           pivots checked), L written back symmetric;
//   2. k_eig_tridiag        one workgroup per patch: C = L^{-1} K L^{-T} by
//                           row and column substitution, then Householder
//                           tridiagonalization H_{n-3}..H_0 C H_0..H_{n-3} = T,
//                           H_k = I - v_k v_k^T / h_k; v_k goes to row k of Kp;
//   3. k_eig_ql             one wave per patch: implicit QL with Wilkinson
//                           shifts on T (eigenvalues), logging every Givens
//                           rotation (index, cos, sin) - a chain of scalar
//                           operations, so many patches share a CU;
//   4. k_eig_vectors        one workgroup per patch: Z = I, the logged
//                           rotations applied row by row, then Z <- H_0..H_{n-3} Z
//                           and Z <- L^{-T} Z; eigenvector k to column k of Kp.
// Chunk buffers: Mp, Kp (column-major n x n per patch, symmetric on entry),
// d/e/h (n per patch), the rotation log (cap entries per patch).
#include "kernels.h"

#include <mutex>
#include <vector>

namespace stfem
{
  namespace
  {
    constexpr int EIG_MAXN = 128;
    // One workgroup of EIG_T threads per patch (the patch matrix fills most
    // of a CU's LDS, so a CU holds one patch; the threads of the 2D loops
    // form an EIG_TI x EIG_TJ grid, the matrix-vector products EIG_P
    // partial sums per row)
    constexpr int EIG_T  = 1024;
    constexpr int EIG_TJ = 32;
    constexpr int EIG_TI = EIG_T / EIG_TJ;
    constexpr int EIG_P  = EIG_T / EIG_MAXN;

    __device__ __forceinline__ double wave_sum(double s)
    {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1)
        s += __shfl_xor(s, o, 64);
      return s;
    }

    // M = L L^T of the patch mass matrix; on exit Mp holds L in both
    // triangles (Mp[i + j n] = L[max(i,j)][min(i,j)]).  info = 1 for a
    // non-positive pivot.
    __global__ void __launch_bounds__(EIG_T)
      k_eig_cholesky(int n, double *__restrict__ Mp, int *__restrict__ info)
    {
      extern __shared__ double A[]; // n x n, row-major
      const size_t base = size_t(blockIdx.x) * n * n;
      const int    tid = threadIdx.x, ti = tid / EIG_TJ, tj = tid % EIG_TJ;
      for (int e = tid; e < n * n; e += EIG_T)
        A[e] = Mp[base + e];
      // right-looking elimination of the upper triangle: after step k, row
      // k holds d_k L[j][k] (j > k) with the pivot d_k = L[k][k]^2
      for (int k = 0; k < n; ++k)
        {
          __syncthreads();
          const double dk = A[k * n + k];
          if (!(dk > 0.0))
            {
              if (tid == 0)
                info[blockIdx.x] = 1;
              return; // uniform: every thread read the same pivot
            }
          const double inv = 1.0 / dk;
          for (int i = k + 1 + ti; i < n; i += EIG_TI)
            {
              const double f = A[k * n + i] * inv;
              for (int j = k + 1 + tj; j < n; j += EIG_TJ)
                if (j >= i)
                  A[i * n + j] -= f * A[k * n + j];
            }
        }
      __syncthreads();
      for (int e = tid; e < n * n; e += EIG_T)
        {
          const int j = e / n, i = e - j * n;
          const int lo = i < j ? i : j, hi = i < j ? j : i;
          Mp[base + e] = A[lo * n + hi] / sqrt(A[lo * n + lo]);
        }
      if (tid == 0)
        info[blockIdx.x] = 0;
    }

    // C = L^{-1} K L^{-T}, then Householder tridiagonalization of C:
    // d[k] = T[k][k], e[k] = T[k][k+1] (e[n-1] = 0), h[k] and v_k in row k
    // of Kp (columns k+1 .. n-1); h[k] = 0 marks an identity step.
    __global__ void __launch_bounds__(EIG_T)
      k_eig_tridiag(int n, const double *__restrict__ Lm, double *__restrict__ Kp,
                    double *__restrict__ dd, double *__restrict__ ee, double *__restrict__ hh)
    {
      extern __shared__ double A[]; // n x n row-major, v[128], p[EIG_P][128], lc[2][128]
      double *const v  = A + n * n;
      double *const p  = v + EIG_MAXN;
      double *const lc = p + EIG_P * EIG_MAXN; // columns k, k+1 of L
      __shared__ double sc[3];
      const size_t base = size_t(blockIdx.x) * n * n, vb = size_t(blockIdx.x) * n;
      const double *L   = Lm + base; // L(i, k) = L[i + k n] (symmetric storage)
      const int    tid = threadIdx.x, ti = tid / EIG_TJ, tj = tid % EIG_TJ;
      for (int e = tid; e < n * n; e += EIG_T)
        A[e] = Kp[base + e];
      if (tid < n)
        lc[tid] = L[tid];
      // rows: X = L^{-1} K (row k unscaled until the end); column k + 1 of
      // L is fetched while step k runs
      for (int k = 0; k < n; ++k)
        {
          __syncthreads();
          const double *lk  = lc + (k & 1) * EIG_MAXN;
          const bool    nxt = k + 1 < n && tid < n;
          const double  ln  = nxt ? L[tid + (k + 1) * n] : 0.0;
          const double  rk  = 1.0 / lk[k];
          for (int i = k + 1 + ti; i < n; i += EIG_TI)
            {
              const double f = lk[i] * rk;
              for (int j = tj; j < n; j += EIG_TJ)
                A[i * n + j] -= f * A[k * n + j];
            }
          if (nxt)
            lc[((k + 1) & 1) * EIG_MAXN + tid] = ln;
        }
      __syncthreads();
      for (int i = ti; i < n; i += EIG_TI)
        {
          const double rk = 1.0 / L[i + i * n];
          for (int j = tj; j < n; j += EIG_TJ)
            A[i * n + j] *= rk;
        }
      __syncthreads();
      if (tid < n)
        lc[tid] = L[tid];
      // columns: C = X L^{-T} (column k unscaled until the end)
      for (int k = 0; k < n; ++k)
        {
          __syncthreads();
          const double *lk  = lc + (k & 1) * EIG_MAXN;
          const bool    nxt = k + 1 < n && tid < n;
          const double  ln  = nxt ? L[tid + (k + 1) * n] : 0.0;
          const double  rk  = 1.0 / lk[k];
          for (int j = k + 1 + tj; j < n; j += EIG_TJ)
            {
              const double f = lk[j] * rk;
              for (int i = ti; i < n; i += EIG_TI)
                A[i * n + j] -= A[i * n + k] * f;
            }
          if (nxt)
            lc[((k + 1) & 1) * EIG_MAXN + tid] = ln;
        }
      __syncthreads();
      for (int j = tj; j < n; j += EIG_TJ)
        {
          const double rk = 1.0 / L[j + j * n];
          for (int i = ti; i < n; i += EIG_TI)
            A[i * n + j] *= rk;
        }
      __syncthreads();
      // exact symmetry
      for (int i = ti; i < n; i += EIG_TI)
        for (int j = tj; j < n; j += EIG_TJ)
          if (j > i)
            {
              const double s = 0.5 * (A[i * n + j] + A[j * n + i]);
              A[i * n + j]   = s;
              A[j * n + i]   = s;
            }
      // Householder steps k = 0 .. n-3 on the trailing block (order m)
      for (int k = 0; k + 2 < n; ++k)
        {
          const int m = n - k - 1, o = k + 1;
          __syncthreads();
          if (tid < 64)
            {
              double s = 0;
              for (int j = tid; j < m; j += 64)
                {
                  const double x = A[k * n + o + j];
                  s += x * x;
                }
              s = wave_sum(s);
              if (tid == 0)
                {
                  const double x0 = A[k * n + o];
                  double       al = x0, h = 0;
                  if (s > 0)
                    {
                      al = -copysign(sqrt(s), x0);
                      h  = s - al * x0;
                    }
                  sc[0] = h;
                  sc[1] = al;
                  sc[2] = x0 - al;
                  dd[vb + k] = A[k * n + k];
                  ee[vb + k] = al;
                  hh[vb + k] = h;
                }
            }
          __syncthreads();
          const double h = sc[0];
          if (h == 0.0)
            continue; // uniform
          for (int j = tid; j < m; j += EIG_T)
            {
              const double vj = j ? A[k * n + o + j] : sc[2];
              v[j]            = vj;
              Kp[base + size_t(k) * n + o + j] = vj;
            }
          __syncthreads();
          // p = C' v / h, EIG_P partial sums per row (column reads: C'
          // symmetric)
          {
            const int i = tid % EIG_MAXN, part = tid / EIG_MAXN;
            if (i < m)
              {
                double s = 0;
                for (int j = part; j < m; j += EIG_P)
                  s += A[(o + j) * n + o + i] * v[j];
                p[part * EIG_MAXN + i] = s;
              }
          }
          __syncthreads();
          // w = p - (v^T p / 2h) v, into p[0][*]
          if (tid < 64)
            {
              double s = 0;
              for (int i = tid; i < m; i += 64)
                {
                  double pi = p[i];
#pragma unroll
                  for (int q = 1; q < EIG_P; ++q)
                    pi += p[q * EIG_MAXN + i];
                  pi /= h;
                  p[i]            = pi;
                  s += pi * v[i];
                }
              const double kc = wave_sum(s) / (2.0 * h);
              for (int i = tid; i < m; i += 64)
                p[i] -= kc * v[i];
            }
          __syncthreads();
          for (int i = ti; i < m; i += EIG_TI)
            {
              const double vi = v[i], wi = p[i];
              for (int j = tj; j < m; j += EIG_TJ)
                A[(o + i) * n + o + j] -= vi * p[j] + wi * v[j];
            }
        }
      __syncthreads();
      if (tid == 0)
        {
          if (n >= 2)
            {
              dd[vb + n - 2] = A[(n - 2) * n + n - 2];
              ee[vb + n - 2] = A[(n - 2) * n + n - 1];
            }
          dd[vb + n - 1] = A[(n - 1) * n + n - 1];
          ee[vb + n - 1] = 0.0;
        }
    }

    // Implicit QL with Wilkinson shifts on the tridiagonal (d, e) of one
    // patch, one wave per patch: every lane runs the same scalar chain on the
    // wave's copy in LDS.  On exit dd holds the eigenvalues.  Every QL
    // iteration is one chain of rotations of the column pairs (i, i+1),
    // i = top, top-1, ..., top-len+1: its (cos, sin) go to rot in order and
    // (top, len) to chains; cnt[2c] = rotations, cnt[2c+1] = chains of patch
    // c.  info = 2 when an eigenvalue needs more than 60 iterations, 3 when
    // a log would overflow (cap rotations, ccap chains).
    __global__ void __launch_bounds__(64)
      k_eig_ql(int n, double *__restrict__ dd, const double *__restrict__ ee, long cap, int ccap,
               double2 *__restrict__ rot, int2 *__restrict__ chains, int *__restrict__ cnt_out,
               int *__restrict__ info)
    {
      __shared__ double d[EIG_MAXN], e[EIG_MAXN];
      const size_t vb = size_t(blockIdx.x) * n;
      const int    lane = threadIdx.x;
      for (int i = lane; i < n; i += 64)
        {
          d[i] = dd[vb + i];
          e[i] = ee[vb + i];
        }
      __syncthreads();
      double2 *const R   = rot + size_t(blockIdx.x) * cap;
      int2 *const    C   = chains + size_t(blockIdx.x) * ccap;
      const double   eps = 0x1.0p-52;
      long           cnt = 0;
      int            nch = 0, bad = 0;
      for (int l = 0; l < n && !bad; ++l)
        for (int it = 0;; ++it)
          {
            int m = l;
            while (m < n - 1 && fabs(e[m]) > eps * (fabs(d[m]) + fabs(d[m + 1])))
              ++m;
            if (m == l)
              break;
            if (it == 60 || cnt + (m - l) > cap || nch == ccap)
              {
                bad = it == 60 ? 2 : 3;
                break;
              }
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g        = d[m] - d[l] + e[l] / (g + copysign(r, g));
            double s = 1.0, c = 1.0, pp = 0.0;
            bool   early = false;
            const long c0 = cnt;
            for (int i = m - 1; i >= l; --i)
              {
                const double f = s * e[i], b = c * e[i];
                r               = hypot(f, g);
                e[i + 1]        = r;
                if (r == 0.0)
                  {
                    d[i + 1] -= pp;
                    e[m]  = 0.0;
                    early = true;
                    break;
                  }
                s        = f / r;
                c        = g / r;
                g        = d[i + 1] - pp;
                r        = (d[i] - g) * s + 2.0 * c * b;
                pp       = s * r;
                d[i + 1] = g + pp;
                g        = c * r - b;
                if (lane == 0)
                  R[cnt] = make_double2(c, s);
                ++cnt;
              }
            if (cnt > c0)
              {
                if (lane == 0)
                  C[nch] = make_int2(m - 1, int(cnt - c0));
                ++nch;
              }
            if (early)
              continue;
            d[l] -= pp;
            e[l] = g;
            e[m] = 0.0;
          }
      __syncthreads();
      for (int i = lane; i < n; i += 64)
        dd[vb + i] = d[i];
      if (lane == 0)
        {
          cnt_out[2 * blockIdx.x]     = int(cnt);
          cnt_out[2 * blockIdx.x + 1] = nch;
          if (bad)
            info[blockIdx.x] = bad;
        }
    }

    // Eigenvectors: Z = I with the logged rotations applied (rows r = tid),
    // Z <- H_0 .. H_{n-3} Z, Z <- L^{-T} Z; column c of Z (eigenvalue dd[c])
    // to column c of Kp.
    //
    // Rotations, chain by chain: the value of the chain's lower column is
    // carried in a register (zc) while the chain walks down, so each rotation
    // reads one column and writes one; the (cos, sin) are held 64 per lane
    // register (this block and the next one in flight) and broadcast with
    // readlane.
    __global__ void __launch_bounds__(EIG_T)
      k_eig_vectors(int n, const double *__restrict__ Lm, double *__restrict__ Kp,
                    const double *__restrict__ hh, long cap, int ccap,
                    const double2 *__restrict__ rot, const int2 *__restrict__ chains,
                    const int *__restrict__ cnt_in)
    {
      // n x n row-major, v[2][128], w[EIG_P][128], lc[2][128], hs[128]
      extern __shared__ double Z[];
      double *const v  = Z + n * n;
      double *const w  = v + 2 * EIG_MAXN;
      double *const lc = w + EIG_P * EIG_MAXN;
      double *const hs = lc + 2 * EIG_MAXN;
      const size_t  base = size_t(blockIdx.x) * n * n, vb = size_t(blockIdx.x) * n;
      const double *L    = Lm + base;
      const int     tid = threadIdx.x, ti = tid / EIG_TJ, tj = tid % EIG_TJ;
      for (int e = tid; e < n * n; e += EIG_T)
        Z[e] = (e % (n + 1) == 0) ? 1.0 : 0.0;
      for (int k = tid; k + 2 < n; k += EIG_T)
        hs[k] = hh[vb + k];
      if (n >= 3 && tid < 2)
        v[((n - 3) & 1) * EIG_MAXN + tid] = Kp[base + size_t(n - 3) * n + n - 2 + tid];
      __syncthreads();
      if (tid < ((n + 63) & ~63))
        {
          const double2 *R    = rot + size_t(blockIdx.x) * cap;
          const int2    *C    = chains + size_t(blockIdx.x) * ccap;
          const int      nr   = cnt_in[2 * blockIdx.x], nch = cnt_in[2 * blockIdx.x + 1];
          const int      lane = tid & 63;
          // lanes past the last row work on a dummy row (w, unused until the
          // Householder steps): no lane masking inside the chains
          double *const zr = tid < n ? Z + tid * n : w;
          // (cos, sin) of rotations [tb, tb + 128): b0 one per lane, b1 the
          // next 64; the window moves by 64 once a group starts past tb + 64
          auto fetch = [&](int t) { return t < nr ? R[t] : make_double2(1.0, 0.0); };
          double2 b0 = fetch(lane), b1 = fetch(64 + lane);
          int     tb = 0, t = 0;
          auto rl = [](double x, int q) {
            return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), q),
                                    __builtin_amdgcn_readlane(__double2loint(x), q));
          };
          auto cs_of = [&](int q, double &c, double &sn) {
            if (q < 64)
              {
                c  = rl(b0.x, q);
                sn = rl(b0.y, q);
              }
            else
              {
                c  = rl(b1.x, q - 64);
                sn = rl(b1.y, q - 64);
              }
          };
          for (int h = 0; h < nch; ++h)
            {
              const int2    hd  = C[h];
              const int     top = hd.x, len = hd.y;
              double *const zt  = zr + top; // zt[-j] = column top - j
              double        zc  = zt[1];
              int           j   = 0;
              for (; j + 8 <= len; j += 8, t += 8)
                {
                  if (t - tb >= 64)
                    {
                      tb += 64;
                      b0 = b1;
                      b1 = fetch(tb + 64 + lane);
                    }
                  double lo[8];
#pragma unroll
                  for (int k = 0; k < 8; ++k)
                    lo[k] = zt[-(j + k)];
                  const int q0 = t - tb;
                  if (q0 + 8 <= 64)
                    {
                      // the common case: the whole group in b0
#pragma unroll
                      for (int k = 0; k < 8; ++k)
                        {
                          const double c = rl(b0.x, q0 + k), sn = rl(b0.y, q0 + k);
                          zt[1 - (j + k)] = sn * lo[k] + c * zc;
                          zc              = c * lo[k] - sn * zc;
                        }
                    }
                  else
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                      {
                        double c, sn;
                        cs_of(q0 + k, c, sn);
                        zt[1 - (j + k)] = sn * lo[k] + c * zc;
                        zc              = c * lo[k] - sn * zc;
                      }
                }
              for (; j < len; ++j, ++t)
                {
                  if (t - tb >= 64)
                    {
                      tb += 64;
                      b0 = b1;
                      b1 = fetch(tb + 64 + lane);
                    }
                  double c, sn;
                  cs_of(t - tb, c, sn);
                  const double lo = zt[-j];
                  zt[1 - j]       = sn * lo + c * zc;
                  zc              = c * lo - sn * zc;
                }
              zt[1 - len] = zc;
            }
        }
      // Z <- H_k Z for k = n-3 .. 0; v_{k-1} fetched while step k runs
      for (int k = n - 3; k >= 0; --k)
        {
          __syncthreads();
          const int     m = n - k - 1, o = k + 1;
          const double  h = hs[k];
          const double *vk = v + (k & 1) * EIG_MAXN;
          const bool    nx = k >= 1 && tid < m + 1;
          const double  vn = nx ? Kp[base + size_t(k - 1) * n + k + tid] : 0.0;
          if (h != 0.0) // uniform
            {
              {
                const int c = tid % EIG_MAXN, part = tid / EIG_MAXN;
                if (c < n)
                  {
                    double s = 0;
                    for (int i = part; i < m; i += EIG_P)
                      s += vk[i] * Z[(o + i) * n + c];
                    w[part * EIG_MAXN + c] = s;
                  }
              }
              __syncthreads();
              if (tid < n)
                {
                  double s = w[tid];
#pragma unroll
                  for (int q = 1; q < EIG_P; ++q)
                    s += w[q * EIG_MAXN + tid];
                  w[tid] = s;
                }
              __syncthreads();
              for (int i = ti; i < m; i += EIG_TI)
                {
                  const double vi = vk[i] / h;
                  for (int c = tj; c < n; c += EIG_TJ)
                    Z[(o + i) * n + c] -= vi * w[c];
                }
            }
          if (nx)
            v[((k - 1) & 1) * EIG_MAXN + tid] = vn;
        }
      // Z <- L^{-T} Z: back substitution, row k scaled at the end; column
      // k - 1 of L (= row k - 1 below the diagonal) fetched during step k
      __syncthreads();
      if (tid < n)
        lc[((n - 1) & 1) * EIG_MAXN + tid] = L[tid + (n - 1) * n];
      for (int k = n - 1; k >= 1; --k)
        {
          __syncthreads();
          const double *lk  = lc + (k & 1) * EIG_MAXN;
          const bool    nxt = k - 1 >= 1 && tid < n;
          const double  ln  = nxt ? L[tid + (k - 1) * n] : 0.0;
          const double  rk  = 1.0 / lk[k];
          for (int i = ti; i < k; i += EIG_TI)
            {
              const double f = lk[i] * rk; // L[k][i]
              for (int c = tj; c < n; c += EIG_TJ)
                Z[i * n + c] -= f * Z[k * n + c];
            }
          if (nxt)
            lc[((k - 1) & 1) * EIG_MAXN + tid] = ln;
        }
      __syncthreads();
      for (int e = tid; e < n * n; e += EIG_T)
        {
          const int c = e / n, r = e - c * n; // column-major output position
          Kp[base + e] = Z[r * n + c] / L[r + r * n];
        }
    }
  } // namespace

  size_t patch_eigen_lds(int n)
  {
    return sizeof(double) * (size_t(n) * n + (5 + EIG_P) * EIG_MAXN);
  }

  // implicit QL needs about n^2 rotations per patch (0.7 - 1.3 n^2 measured
  // on random and on patch matrices); a chunk with a patch beyond the cap is
  // redone by the library eigensolver
  long patch_eigen_rotation_cap(int n)
  {
    return 2L * n * n + 256;
  }
  // QL iterations (one chain each) per patch: about n measured
  int patch_eigen_chain_cap(int n)
  {
    return 8 * n + 64;
  }

  void launch_patch_eigen(int n, long np, double *Mp, double *Kp, double *lam, double *ee,
                          double *hh, double2 *rot, int2 *chains, int *counts, int *info,
                          hipStream_t s)
  {
    STFEM_REQUIRE(n >= 1 && n <= EIG_MAXN, "patch eigensolver: at most 128 nodes per patch");
    if (np == 0)
      return;
    const size_t lds = patch_eigen_lds(n);
    {
      // LDS beyond 64 KiB per workgroup, once per device
      static std::mutex        mtx;
      static std::vector<char> set;
      int                      dev = 0;
      STFEM_HIP(hipGetDevice(&dev));
      std::lock_guard<std::mutex> lk(mtx);
      if (dev >= int(set.size()))
        set.resize(dev + 1, 0);
      if (!set[dev])
        {
          const int mx = int(patch_eigen_lds(EIG_MAXN));
          STFEM_HIP(hipFuncSetAttribute((const void *)k_eig_cholesky,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, mx));
          STFEM_HIP(hipFuncSetAttribute((const void *)k_eig_tridiag,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, mx));
          STFEM_HIP(hipFuncSetAttribute((const void *)k_eig_vectors,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, mx));
          set[dev] = 1;
        }
    }
    const long cap = patch_eigen_rotation_cap(n);
    const dim3 grid{(unsigned)np, 1, 1};
    k_eig_cholesky<<<grid, EIG_T, sizeof(double) * size_t(n) * n, s>>>(n, Mp, info);
    STFEM_HIP(hipGetLastError());
    k_eig_tridiag<<<grid, EIG_T, lds, s>>>(n, Mp, Kp, lam, ee, hh);
    STFEM_HIP(hipGetLastError());
    const int  ccap = patch_eigen_chain_cap(n);
    k_eig_ql<<<grid, 64, 0, s>>>(n, lam, ee, cap, ccap, rot, chains, counts, info);
    STFEM_HIP(hipGetLastError());
    k_eig_vectors<<<grid, EIG_T, lds, s>>>(n, Mp, Kp, hh, cap, ccap, rot, chains, counts);
    STFEM_HIP(hipGetLastError());
  }
} // namespace stfem
