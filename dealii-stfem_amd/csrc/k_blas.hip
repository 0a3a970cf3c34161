// Vector kernels of the Krylov and multigrid drivers (gfx950).
//
// Reductions are two-stage with a fixed grid so results are bitwise
// reproducible run to run: stage 1 writes one partial sum per workgroup;
// stage 2 sums the partials in a fixed order, either in a one-workgroup
// kernel, or at the start of the kernel that consumes the result (the
// Gram-Schmidt chain: each step re-reduces the previous step's partials, so
// the chain needs no extra launches and no host round trip), or on the host.
#include "kernels.h"

#include <algorithm>
#include <cmath>

namespace stfem
{
  namespace
  {
    constexpr int RB = 1024; // stage-1 workgroups
    constexpr int RT = 256;  // threads per workgroup

    __device__ __forceinline__ double block_sum(double v)
    {
      __shared__ double red[RT / 64];
      for (int off = 32; off > 0; off >>= 1)
        v += __shfl_down(v, off, 64);
      const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
      if (lane == 0)
        red[w] = v;
      __syncthreads();
      double s = 0.0;
      if (threadIdx.x == 0)
        for (int i = 0; i < RT / 64; ++i)
          s += red[i];
      return s;
    }

    template <typename T>
    __global__ void __launch_bounds__(RT)
      k_dot_partial(long n, const T *__restrict__ x, const T *__restrict__ y, double *partial)
    {
      double s = 0.0;
      for (long i = (long)blockIdx.x * RT + threadIdx.x; i < n; i += (long)RB * RT)
        s += double(x[i]) * double(y[i]);
      s = block_sum(s);
      if (threadIdx.x == 0)
        partial[blockIdx.x] = s;
    }

    __global__ void __launch_bounds__(RT) k_finish(const double *partial, double *out, int op)
    {
      double s = 0.0;
      for (int i = threadIdx.x; i < RB; i += RT)
        s += partial[i];
      s = block_sum(s);
      if (threadIdx.x == 0)
        out[0] = (op == 1) ? sqrt(s) : s;
    }

    // x += factor * c * y, c = sum of the previous reduction's partials
    // `pin` (every workgroup sums them in the same order, so all use the
    // same c; workgroup 0 also stores c to *hout), then the partial sums of
    // dot(x, z) (z null: dot(x, x)) into `pout`.
    __global__ void __launch_bounds__(RT)
      k_axpy_dot_chain(long n, double *__restrict__ x, const double *__restrict__ pin,
                       double factor, const double *__restrict__ y,
                       const double *__restrict__ z, double *__restrict__ pout, double *hout)
    {
      __shared__ double cs;
      double v = 0.0;
      for (int i = threadIdx.x; i < RB; i += RT)
        v += pin[i];
      v = block_sum(v);
      if (threadIdx.x == 0)
        {
          cs = v;
          if (blockIdx.x == 0 && hout)
            *hout = v;
        }
      __syncthreads();
      const double c = cs * factor;
      double       s = 0.0;
      for (long i = (long)blockIdx.x * RT + threadIdx.x; i < n; i += (long)RB * RT)
        {
          const double w = x[i] + c * y[i];
          x[i]           = w;
          s += w * (z ? z[i] : w);
        }
      s = block_sum(s);
      if (threadIdx.x == 0)
        pout[blockIdx.x] = s;
    }

    template <typename T>
    __global__ void k_axpby(long n, T a, const T *__restrict__ x, T b, T *__restrict__ y)
    {
      const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (i < n)
        y[i] = a * x[i] + (b == T(0) ? T(0) : b * y[i]);
    }
    // the same with 16 bytes per thread (vectors 16-byte aligned); the last
    // n % VW entries go to the first thread
    template <typename T>
    __global__ void k_axpby_vec(long n, T a, const T *__restrict__ x, T b, T *__restrict__ y)
    {
      constexpr int VW = 16 / int(sizeof(T));
      typedef T     vec_t __attribute__((ext_vector_type(VW)));
      const long    i  = (long)blockIdx.x * blockDim.x + threadIdx.x;
      const long    nv = n / VW;
      if (i < nv)
        {
          const vec_t xv = reinterpret_cast<const vec_t *>(x)[i];
          vec_t       r;
          if (b == T(0))
            r = a * xv;
          else
            {
              const vec_t yv = reinterpret_cast<const vec_t *>(y)[i];
#pragma unroll
              for (int e = 0; e < VW; ++e)
                r[e] = a * xv[e] + b * yv[e];
            }
          reinterpret_cast<vec_t *>(y)[i] = r;
        }
      if (i == 0)
        for (long t = nv * VW; t < n; ++t)
          y[t] = a * x[t] + (b == T(0) ? T(0) : b * y[t]);
    }

    struct MultiAxpy
    {
      const double *z[32];
      double        c[32];
      int           m;
    };
    __global__ void k_multi_axpy(long n, MultiAxpy A, double *__restrict__ x)
    {
      const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (i >= n)
        return;
      double s = x[i];
      for (int j = 0; j < A.m; ++j)
        s += A.c[j] * A.z[j][i];
      x[i] = s;
    }

    // ---- Gram-Schmidt with one reduction pass per Krylov step
    // Stage 1 of the dots (v_i, x) and (v_i, y), i < m, and (x, y): the
    // partial sums of workgroup b go to partial[d * RB + b] with d = dx + i,
    // dy + i and dxy (dxy < 0: no (x, y) dot).
    typedef double vd2 __attribute__((ext_vector_type(2)));
    constexpr int  MD = 16; // left vectors per pass
    struct DotList
    {
      const double *v[MD];
      int           m, dx, dy, dxy;
    };
    // VEC: every vector 16-byte aligned; thread work items are pairs of
    // consecutive entries read as one 16-byte load each (the last entry of
    // an odd n goes to the first thread)
    template <bool VEC>
    __global__ void __launch_bounds__(RT)
      k_multi_dot(long n, DotList L, const double *__restrict__ x, const double *__restrict__ y,
                  double *__restrict__ partial)
    {
      double ax[MD], ay[MD], axy = 0.0;
#pragma unroll
      for (int i = 0; i < MD; ++i)
        ax[i] = ay[i] = 0.0;
      auto one = [&](long e) {
        const double xe = x[e], ye = y[e];
        axy += xe * ye;
#pragma unroll
        for (int i = 0; i < MD; ++i)
          if (i < L.m)
            {
              const double ve = L.v[i][e];
              ax[i] += ve * xe;
              ay[i] += ve * ye;
            }
      };
      if constexpr (VEC)
        {
          const long n2 = n / 2;
          for (long e = (long)blockIdx.x * RT + threadIdx.x; e < n2; e += (long)RB * RT)
            {
              const double2 xe = reinterpret_cast<const double2 *>(x)[e];
              const double2 ye = reinterpret_cast<const double2 *>(y)[e];
              axy += xe.x * ye.x;
              axy += xe.y * ye.y;
#pragma unroll
              for (int i = 0; i < MD; ++i)
                if (i < L.m)
                  {
                    // basis vectors stream through once per pass
                    const vd2 ve = __builtin_nontemporal_load(
                      reinterpret_cast<const vd2 *>(L.v[i]) + e);
                    ax[i] += ve.x * xe.x;
                    ax[i] += ve.y * xe.y;
                    ay[i] += ve.x * ye.x;
                    ay[i] += ve.y * ye.y;
                  }
            }
          if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0)
            one(n - 1);
        }
      else
        for (long e = (long)blockIdx.x * RT + threadIdx.x; e < n; e += (long)RB * RT)
          one(e);
      // wavefront sums, then the four wavefront sums in order
      __shared__ double red[RT / 64][2 * MD + 1];
      const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
      for (int d = 0; d < 2 * MD + 1; ++d)
        {
          if (d < 2 * MD && (d % MD) >= L.m)
            continue;
          double v = d < MD ? ax[d] : (d < 2 * MD ? ay[d - MD] : axy);
          for (int off = 32; off > 0; off >>= 1)
            v += __shfl_down(v, off, 64);
          if (lane == 0)
            red[w][d] = v;
        }
      __syncthreads();
      const int d = threadIdx.x;
      if (d < 2 * MD + 1 && (d == 2 * MD ? L.dxy >= 0 : (d % MD) < L.m))
        {
          double s = 0.0;
          for (int i = 0; i < RT / 64; ++i)
            s += red[i][d];
          const int o = d < MD ? L.dx + d : (d < 2 * MD ? L.dy + d - MD : L.dxy);
          partial[(long)o * RB + blockIdx.x] = s;
        }
    }

    // stage 2 of nd dots (one workgroup each, the order of k_finish)
    __global__ void __launch_bounds__(RT) k_finish_many(const double *partial, double *out)
    {
      const double *p = partial + (long)blockIdx.x * RB;
      double        s = 0.0;
      for (int i = threadIdx.x; i < RB; i += RT)
        s += p[i];
      s = block_sum(s);
      if (threadIdx.x == 0)
        out[blockIdx.x] = s;
    }

    // Coefficients of step j (j < MG) from dots = [(v_0,x) .. (v_j,x),
    // (v_0,v_j) .. (v_{j-1},v_j)]: column j of the Gram matrix G (ld ldg) is
    // stored, then h_i = (v_i, x) - sum_{k<i} G(k,i) h_k, which equals the
    // coefficient modified Gram-Schmidt computes from x - sum_{k<i} h_k v_k.
    // One workgroup of MG threads stages the upper triangle of columns 0..j
    // in LDS; one thread then runs the recurrence from LDS.
    constexpr int MG = 64;
    __global__ void __launch_bounds__(MG)
      k_gs_coefficients(const double *dots, int j, double *G, int ldg, double *h)
    {
      __shared__ double Gs[MG][MG]; // Gs[i][k] = G(k, i), k < i <= j
      __shared__ double hs[MG];
      const int         t = threadIdx.x, n = j + 1;
      for (int i = t; i < j; i += MG)
        G[i + (long)j * ldg] = dots[j + 1 + i];
      for (int e = t; e < n * n; e += MG)
        {
          const int i = e / n, k = e - i * n;
          if (k < i)
            Gs[i][k] = (i == j) ? dots[j + 1 + k] : G[k + (long)i * ldg];
        }
      if (t < n)
        hs[t] = dots[t];
      __syncthreads();
      if (t != 0)
        return;
      for (int i = 0; i <= j; ++i)
        {
          double s = hs[i];
          for (int k = 0; k < i; ++k)
            s -= Gs[i][k] * hs[k];
          hs[i] = s;
          h[i]  = s;
        }
    }

    // x -= sum_i h_i v_i (h on the device), then the partial sums of (x, x)
    constexpr int MA = 64;
    struct AxpyList
    {
      const double *v[MA];
      int           m;
    };
    // VEC: as in k_multi_dot (pairs of entries, 16-byte accesses)
    template <bool VEC>
    __global__ void __launch_bounds__(RT)
      k_multi_axpy_norm(long n, AxpyList L, const double *__restrict__ h, double *__restrict__ x,
                        double *__restrict__ pnorm)
    {
      __shared__ double hs[MA];
      if (threadIdx.x < L.m)
        hs[threadIdx.x] = h[threadIdx.x];
      __syncthreads();
      double s    = 0.0;
      auto   one  = [&](long e) {
        double w = x[e];
        for (int i = 0; i < L.m; ++i)
          w -= hs[i] * L.v[i][e];
        x[e] = w;
        s += w * w;
      };
      if constexpr (VEC)
        {
          const long n2 = n / 2;
          for (long e = (long)blockIdx.x * RT + threadIdx.x; e < n2; e += (long)RB * RT)
            {
              double2 w = reinterpret_cast<const double2 *>(x)[e];
#pragma unroll 4
              for (int i = 0; i < L.m; ++i)
                {
                  const vd2 v = __builtin_nontemporal_load(reinterpret_cast<const vd2 *>(L.v[i]) + e);
                  w.x -= hs[i] * v.x;
                  w.y -= hs[i] * v.y;
                }
              // streamed out (whole 128-byte lines per wavefront), not kept
              // in the caches that the basis vectors stream through
              vd2 wv = {w.x, w.y};
              __builtin_nontemporal_store(wv, reinterpret_cast<vd2 *>(x) + e);
              s += w.x * w.x;
              s += w.y * w.y;
            }
          if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0)
            one(n - 1);
        }
      else
        for (long e = (long)blockIdx.x * RT + threadIdx.x; e < n; e += (long)RB * RT)
          one(e);
      s = block_sum(s);
      if (threadIdx.x == 0)
        pnorm[blockIdx.x] = s;
    }

    template <typename TD, typename TS>
    __global__ void k_convert(long n, const TS *__restrict__ s, TD *__restrict__ d)
    {
      const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (i < n)
        d[i] = TD(s[i]);
    }

    // four entries per thread (16- and 32-byte accesses; vectors 32-byte
    // aligned); the last n % 4 entries go to the first thread
    template <typename TD, typename TS>
    __global__ void k_convert4(long n, const TS *__restrict__ s, TD *__restrict__ d)
    {
      typedef TS vs_t __attribute__((ext_vector_type(4)));
      typedef TD vd_t __attribute__((ext_vector_type(4)));
      const long i  = (long)blockIdx.x * blockDim.x + threadIdx.x;
      const long nv = n / 4;
      if (i < nv)
        {
          const vs_t v = reinterpret_cast<const vs_t *>(s)[i];
          vd_t       r;
#pragma unroll
          for (int e = 0; e < 4; ++e)
            r[e] = TD(v[e]);
          reinterpret_cast<vd_t *>(d)[i] = r;
        }
      if (i == 0)
        for (long t = nv * 4; t < n; ++t)
          d[t] = TD(s[t]);
    }
    template <typename T>
    __global__ void k_b2i(long nn, int nt, const T *__restrict__ b, T *__restrict__ it)
    {
      const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (i >= nn * nt)
        return;
      const long node = i / nt;
      const int  t    = int(i - node * nt);
      it[i]           = b[(long)t * nn + node];
    }
    template <typename T>
    __global__ void k_i2b(long nn, int nt, const T *__restrict__ it, T *__restrict__ b)
    {
      const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (i >= nn * nt)
        return;
      const long node = i / nt;
      const int  t    = int(i - node * nt);
      b[(long)t * nn + node] = it[i];
    }

    template <typename T>
    __global__ void k_zero_dirichlet(GridDesc g, int nt, T *x)
    {
      const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (i >= g.n_nodes)
        return;
      const long gx = i % g.nn[0], r = i / g.nn[0];
      const long gy = r % g.nn[1], gz = r / g.nn[1];
      const int  d  = g.dirichlet;
      bool       c  = (gx == 0 && (d & 1)) || (gx == g.nn[0] - 1 && (d & 2)) ||
               (gy == 0 && (d & 4)) || (gy == g.nn[1] - 1 && (d & 8));
      if (g.dim == 3)
        c = c || (gz == 0 && (d & 16)) || (gz == g.nn[2] - 1 && (d & 32));
      if (c)
        for (int t = 0; t < nt; ++t)
          x[i * nt + t] = T(0);
    }

    template <typename T>
    __global__ void k_fill_zero(T *p, size_t n)
    {
      for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
           i += (size_t)gridDim.x * blockDim.x)
        p[i] = T(0);
    }

    inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

    template <typename C>
    __global__ void k_add_diagonal(long n, C a, C *M)
    {
      const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (i < n)
        M[i + i * n] += a;
    }
    template <typename C>
    __global__ void k_scale(long n, C a, C *x)
    {
      const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (i < n)
        x[i] *= a;
    }
    struct ProbeColour
    {
      int c[3];
    };
    __device__ __forceinline__ void lattice(const GridDesc &g, long node, long q[3])
    {
      q[0] = node % g.nn[0];
      const long r = node / g.nn[0];
      q[1]         = g.dim == 3 ? r % g.nn[1] : r;
      q[2]         = g.dim == 3 ? r / g.nn[1] : 0;
    }
    template <typename T>
    __global__ void k_probe_vector(const GridDesc in, int nt_in, ProbeColour c, int t, int stride,
                                   T *__restrict__ x)
    {
      const long n = in.n_nodes * nt_in;
      for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
           i += (long)gridDim.x * blockDim.x)
        {
          const long node = i / nt_in;
          long       q[3];
          lattice(in, node, q);
          bool hit = int(i - node * nt_in) == t;
          for (int d = 0; d < in.dim; ++d)
            hit = hit && q[d] % stride == c.c[d];
          x[i] = hit ? T(1) : T(0);
        }
    }
    template <typename T, typename C>
    __global__ void k_probe_store(const GridDesc in, int nt_in, const GridDesc out, int nt_out,
                                  ProbeColour c, int t, int stride, int sn, int sd, int radius,
                                  const T *__restrict__ y, C *__restrict__ M, long ld,
                                  int *__restrict__ bad)
    {
      const long n = out.n_nodes * nt_out;
      for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
           i += (long)gridDim.x * blockDim.x)
        {
          const T v = y[i];
          long    o[3];
          lattice(out, i / nt_out, o);
          long node = 0, mul = 1;
          bool ok   = true;
          for (int d = 0; d < in.dim && ok; ++d)
            {
              // the colour's inputs g = c + stride m near o: |sn g - sd o| <= sd radius
              const long base = (sd * o[d]) / sn; // g closest from below (scaled)
              long       m0   = (base - c.c[d]) / stride - 1;
              long       hitg = -1;
              for (long m = m0; m <= m0 + 3; ++m)
                {
                  const long g = c.c[d] + stride * m;
                  if (g < 0 || g >= in.nn[d])
                    continue;
                  const long dist = sn * g - sd * o[d];
                  if (dist <= long(sd) * radius && -dist <= long(sd) * radius)
                    hitg = g;
                }
              ok = hitg >= 0;
              node += hitg * mul;
              mul *= in.nn[d];
            }
          if (ok)
            M[(node * nt_in + t) * ld + i] = C(v);
          else if (v != T(0))
            atomicAdd(bad, 1);
        }
    }
    template <typename C>
    __global__ void k_transpose(long m, long n, const C *__restrict__ A, C *__restrict__ B)
    {
      __shared__ C tile[32][33];
      const long i0 = long(blockIdx.x) * 32, j0 = long(blockIdx.y) * 32;
      const int  tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 32 x 8
      for (int k = ty; k < 32; k += 8)
        {
          const long i = i0 + tx, j = j0 + k; // A(i, j), i < m, j < n
          tile[k][tx] = (i < m && j < n) ? A[i + j * m] : C(0);
        }
      __syncthreads();
      for (int k = ty; k < 32; k += 8)
        {
          const long j = j0 + tx, i = i0 + k; // B(j, i) = A(i, j)
          if (i < m && j < n)
            B[j + i * n] = tile[tx][k];
        }
    }
    template <typename T, typename C>
    __global__ void k_pack_rows(long n, const C *__restrict__ M, T *__restrict__ R, long ld)
    {
      const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
      if (idx >= n * ld)
        return;
      const long i = idx / ld, j = idx - i * ld;
      R[idx] = (j < n) ? T(M[i + j * n]) : T(0);
    }

    // y = R x, R row-major with rows of ld entries: one wavefront per row
    // streams the row in 16-byte vectors (coalesced), sums in FP64 and
    // reduces across the wavefront.
    template <typename T>
    __global__ void __launch_bounds__(256)
      k_gemv_rows(long n, long ld, const T *__restrict__ R, const T *__restrict__ x,
                  T *__restrict__ y)
    {
      constexpr int VW = 16 / sizeof(T);
      typedef T     vec_t __attribute__((ext_vector_type(VW)));
      const long    row  = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
      const int     lane = threadIdx.x & 63;
      if (row >= n)
        return; // uniform per wavefront
      const T *r   = R + row * ld;
      double   acc = 0.0;
      const long nv = (n / VW) * VW;
      // the matrix streams through once per application (non-temporal
      // loads); the x chunks stay cached
#pragma unroll 4
      for (long j = (long)lane * VW; j < nv; j += 64 * VW)
        {
          const vec_t m  = __builtin_nontemporal_load(reinterpret_cast<const vec_t *>(r + j));
          const vec_t xv = *reinterpret_cast<const vec_t *>(x + j);
#pragma unroll
          for (int e = 0; e < VW; ++e)
            acc += double(m[e]) * double(xv[e]);
        }
      for (long j = nv + lane; j < n; j += 64)
        acc += double(r[j]) * double(x[j]);
      for (int off = 32; off > 0; off >>= 1)
        acc += __shfl_down(acc, off, 64);
      if (lane == 0)
        y[row] = T(acc);
    }
  } // namespace

  template <typename T>
  void launch_fill_zero(T *p, size_t n, hipStream_t s)
  {
    if (n == 0)
      return;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    k_fill_zero<T><<<dim3((unsigned)blocks), 256, 0, s>>>(p, n);
    STFEM_HIP(hipGetLastError());
  }
  template void launch_fill_zero<double>(double *, size_t, hipStream_t);
  template void launch_fill_zero<float>(float *, size_t, hipStream_t);
  template void launch_fill_zero<int>(int *, size_t, hipStream_t);

  int reduce_slots() { return RB; }

  template <typename T>
  void launch_dot(long n, const T *x, const T *y, double *partial, double *out, int op,
                  hipStream_t s)
  {
    k_dot_partial<T><<<RB, RT, 0, s>>>(n, x, y, partial);
    k_finish<<<1, RT, 0, s>>>(partial, out, op);
    STFEM_HIP(hipGetLastError());
  }

  template <typename T>
  void launch_dot_partials(long n, const T *x, const T *y, double *partial, hipStream_t s)
  {
    k_dot_partial<T><<<RB, RT, 0, s>>>(n, x, y, partial);
    STFEM_HIP(hipGetLastError());
  }

  void launch_axpy_dot_chain(long n, double *x, const double *pin, double factor, const double *y,
                             const double *z, double *pout, double *hout, hipStream_t s)
  {
    k_axpy_dot_chain<<<RB, RT, 0, s>>>(n, x, pin, factor, y, z, pout, hout);
    STFEM_HIP(hipGetLastError());
  }

  double sum_partials(const double *h)
  {
    // the stage-2 order of k_finish: thread t sums t, t+RT, ...; then the
    // wavefront tree and the four wavefront sums
    double part[RT];
    for (int t = 0; t < RT; ++t)
      {
        double v = 0.0;
        for (int i = t; i < RB; i += RT)
          v += h[i];
        part[t] = v;
      }
    double w[RT / 64];
    for (int g = 0; g < RT / 64; ++g)
      {
        double *p = part + 64 * g;
        for (int off = 32; off > 0; off >>= 1)
          for (int l = 0; l < off; ++l)
            p[l] += p[l + off];
        w[g] = p[0];
      }
    double r = 0.0;
    for (int g = 0; g < RT / 64; ++g)
      r += w[g];
    return r;
  }

  void launch_gs_step(long n, int j, const double *const *v, double *x, double *G, int ldg,
                      double *partial, double *dots, double *h, double *pnorm, hipStream_t s)
  {
    STFEM_REQUIRE(j >= 0 && j < MA && j < MG && j < ldg, "Gram-Schmidt step: basis too large");
    const double *vj = v[j];
    // 16-byte accesses when every vector allows them
    auto al  = [](const double *p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
    bool vec = al(x);
    for (int i = 0; i <= j; ++i)
      vec = vec && al(v[i]);
    // (v_i, x) -> d = i, (v_i, v_j) -> d = j + 1 + i for i < j; (v_j, x) -> d = j
    for (int i0 = 0; i0 < j || i0 == 0; i0 += MD)
      {
        DotList L;
        L.m   = std::min(MD, j - i0);
        L.dx  = i0;
        L.dy  = j + 1 + i0;
        L.dxy = (i0 == 0) ? j : -1;
        for (int i = 0; i < L.m; ++i)
          L.v[i] = v[i0 + i];
        if (vec)
          k_multi_dot<true><<<RB, RT, 0, s>>>(n, L, x, vj, partial);
        else
          k_multi_dot<false><<<RB, RT, 0, s>>>(n, L, x, vj, partial);
        if (j == 0)
          break;
      }
    k_finish_many<<<2 * j + 1, RT, 0, s>>>(partial, dots);
    k_gs_coefficients<<<1, MG, 0, s>>>(dots, j, G, ldg, h);
    AxpyList A;
    A.m = j + 1;
    for (int i = 0; i <= j; ++i)
      A.v[i] = v[i];
    if (vec)
      k_multi_axpy_norm<true><<<RB, RT, 0, s>>>(n, A, h, x, pnorm);
    else
      k_multi_axpy_norm<false><<<RB, RT, 0, s>>>(n, A, h, x, pnorm);
    STFEM_HIP(hipGetLastError());
  }

  template <typename T>
  void launch_axpby(long n, T a, const T *x, T b, T *y, hipStream_t s)
  {
    if (n == 0)
      return;
    if (reinterpret_cast<uintptr_t>(x) % 16 == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0)
      k_axpby_vec<T><<<nblk((n + 16 / sizeof(T) - 1) / (16 / sizeof(T))), 256, 0, s>>>(n, a, x, b,
                                                                                     y);
    else
      k_axpby<T><<<nblk(n), 256, 0, s>>>(n, a, x, b, y);
    STFEM_HIP(hipGetLastError());
  }

  void launch_multi_axpy(long n, int m, const double *const *z, const double *coef, double *x,
                         hipStream_t s)
  {
    for (int j0 = 0; j0 < m; j0 += 32)
      {
        MultiAxpy A;
        A.m = (m - j0 < 32) ? m - j0 : 32;
        for (int j = 0; j < A.m; ++j)
          {
            A.z[j] = z[j0 + j];
            A.c[j] = coef[j0 + j];
          }
        k_multi_axpy<<<nblk(n), 256, 0, s>>>(n, A, x);
      }
    STFEM_HIP(hipGetLastError());
  }

  template <typename TD, typename TS>
  void launch_convert(long n, const TS *src, TD *dst, hipStream_t s)
  {
    if (n == 0)
      return;
    if (reinterpret_cast<uintptr_t>(src) % 32 == 0 && reinterpret_cast<uintptr_t>(dst) % 32 == 0)
      k_convert4<TD, TS><<<nblk((n + 3) / 4), 256, 0, s>>>(n, src, dst);
    else
      k_convert<TD, TS><<<nblk(n), 256, 0, s>>>(n, src, dst);
    STFEM_HIP(hipGetLastError());
  }

  template <typename T>
  void launch_blocks_to_interleaved(long nn, int nt, const T *b, T *it, hipStream_t s)
  {
    if (nn * nt == 0)
      return;
    k_b2i<T><<<nblk(nn * nt), 256, 0, s>>>(nn, nt, b, it);
    STFEM_HIP(hipGetLastError());
  }
  template <typename T>
  void launch_interleaved_to_blocks(long nn, int nt, const T *it, T *b, hipStream_t s)
  {
    if (nn * nt == 0)
      return;
    k_i2b<T><<<nblk(nn * nt), 256, 0, s>>>(nn, nt, it, b);
    STFEM_HIP(hipGetLastError());
  }
  template <typename T>
  void launch_zero_dirichlet(const GridDesc &g, int nt, T *x, hipStream_t s)
  {
    if (g.n_nodes == 0)
      return;
    k_zero_dirichlet<T><<<nblk(g.n_nodes), 256, 0, s>>>(g, nt, x);
    STFEM_HIP(hipGetLastError());
  }

  template <typename C>
  void launch_add_diagonal(long n, C a, C *M, hipStream_t s)
  {
    if (n == 0)
      return;
    k_add_diagonal<C><<<nblk(n), 256, 0, s>>>(n, a, M);
    STFEM_HIP(hipGetLastError());
  }
  template <typename C>
  void launch_scale(long n, C a, C *x, hipStream_t s)
  {
    if (n == 0)
      return;
    k_scale<C><<<nblk(n), 256, 0, s>>>(n, a, x);
    STFEM_HIP(hipGetLastError());
  }
  template <typename T, typename C>
  void launch_pack_rows(long n, const C *M, T *R, long ld, hipStream_t s)
  {
    STFEM_REQUIRE(ld >= n, "pack_rows: leading dimension below the order");
    if (n == 0)
      return;
    k_pack_rows<T, C><<<nblk(n * ld), 256, 0, s>>>(n, M, R, ld);
    STFEM_HIP(hipGetLastError());
  }
  template <typename T>
  void launch_gemv_rows(long n, long ld, const T *R, const T *x, T *y, hipStream_t s)
  {
    STFEM_REQUIRE(ld >= n && ld % long(16 / sizeof(T)) == 0,
                  "gemv_rows: rows must be padded to 16-byte multiples");
    if (n == 0)
      return;
    k_gemv_rows<T><<<dim3((unsigned)((n + 3) / 4)), 256, 0, s>>>(n, ld, R, x, y);
    STFEM_HIP(hipGetLastError());
  }

  template <typename T>
  void launch_probe_vector(const GridDesc &in, int nt_in, const int *c, int t, int stride, T *x,
                           hipStream_t s)
  {
    const long n = in.n_nodes * nt_in;
    if (n == 0)
      return;
    ProbeColour pc{{c[0], in.dim > 1 ? c[1] : 0, in.dim > 2 ? c[2] : 0}};
    const long  blocks = std::min<long>((n + 255) / 256, 4096);
    k_probe_vector<T><<<dim3((unsigned)blocks), 256, 0, s>>>(in, nt_in, pc, t, stride, x);
    STFEM_HIP(hipGetLastError());
  }
  template <typename T, typename C>
  void launch_probe_store(const GridDesc &in, int nt_in, const GridDesc &out, int nt_out,
                          const int *c, int t, int stride, int scale_num, int scale_den,
                          int radius, const T *y, C *M, long ld, int *bad, hipStream_t s)
  {
    STFEM_REQUIRE(in.dim == out.dim && stride >= 1 && scale_num >= 1 && scale_den >= 1 &&
                    radius >= 0 && long(stride) * scale_num > 2L * scale_den * radius,
                  "probe: supports of one colour must be disjoint");
    const long n = out.n_nodes * nt_out;
    if (n == 0)
      return;
    ProbeColour pc{{c[0], in.dim > 1 ? c[1] : 0, in.dim > 2 ? c[2] : 0}};
    const long  blocks = std::min<long>((n + 255) / 256, 4096);
    k_probe_store<T, C><<<dim3((unsigned)blocks), 256, 0, s>>>(in, nt_in, out, nt_out, pc, t, stride,
                                                            scale_num, scale_den, radius, y, M, ld,
                                                            bad);
    STFEM_HIP(hipGetLastError());
  }
  template <typename C>
  void launch_transpose(long m, long n, const C *A, C *B, hipStream_t s)
  {
    if (m == 0 || n == 0)
      return;
    const dim3 grid{(unsigned)((m + 31) / 32), (unsigned)((n + 31) / 32), 1};
    k_transpose<C><<<grid, 256, 0, s>>>(m, n, A, B);
    STFEM_HIP(hipGetLastError());
  }
  template void launch_probe_vector<double>(const GridDesc &, int, const int *, int, int, double *,
                                            hipStream_t);
  template void launch_probe_vector<float>(const GridDesc &, int, const int *, int, int, float *,
                                           hipStream_t);
  template void launch_probe_store<double, double>(const GridDesc &, int, const GridDesc &, int,
                                                   const int *, int, int, int, int, int,
                                                   const double *, double *, long, int *,
                                                   hipStream_t);
  template void launch_probe_store<float, float>(const GridDesc &, int, const GridDesc &, int,
                                                 const int *, int, int, int, int, int,
                                                 const float *, float *, long, int *, hipStream_t);
  template void launch_add_diagonal<double>(long, double, double *, hipStream_t);
  template void launch_add_diagonal<float>(long, float, float *, hipStream_t);
  template void launch_scale<double>(long, double, double *, hipStream_t);
  template void launch_scale<float>(long, float, float *, hipStream_t);
  template void launch_transpose<double>(long, long, const double *, double *, hipStream_t);
  template void launch_transpose<float>(long, long, const float *, float *, hipStream_t);
  template void launch_pack_rows<double, double>(long, const double *, double *, long, hipStream_t);
  template void launch_pack_rows<float, float>(long, const float *, float *, long, hipStream_t);
  template void launch_gemv_rows<double>(long, long, const double *, const double *, double *,
                                         hipStream_t);
  template void launch_gemv_rows<float>(long, long, const float *, const float *, float *,
                                        hipStream_t);
  template void launch_dot<double>(long, const double *, const double *, double *, double *, int,
                                   hipStream_t);
  template void launch_dot_partials<double>(long, const double *, const double *, double *,
                                            hipStream_t);
  template void launch_dot_partials<float>(long, const float *, const float *, double *,
                                           hipStream_t);
  template void launch_dot<float>(long, const float *, const float *, double *, double *, int,
                                  hipStream_t);
  template void launch_axpby<double>(long, double, const double *, double, double *, hipStream_t);
  template void launch_axpby<float>(long, float, const float *, float, float *, hipStream_t);
  template void launch_convert<float, double>(long, const double *, float *, hipStream_t);
  template void launch_convert<double, float>(long, const float *, double *, hipStream_t);
  template void launch_blocks_to_interleaved<double>(long, int, const double *, double *,
                                                     hipStream_t);
  template void launch_blocks_to_interleaved<float>(long, int, const float *, float *,
                                                    hipStream_t);
  template void launch_interleaved_to_blocks<double>(long, int, const double *, double *,
                                                     hipStream_t);
  template void launch_interleaved_to_blocks<float>(long, int, const float *, float *,
                                                    hipStream_t);
  template void launch_zero_dirichlet<double>(const GridDesc &, int, double *, hipStream_t);
  template void launch_zero_dirichlet<float>(const GridDesc &, int, float *, hipStream_t);
} // namespace stfem
