// Host side of the MI355X space-time multigrid slab solver: setup (mesh
// hierarchy, element and patch matrices, Vanka inverses, transfer
// operators, relaxation estimates) and the drivers (V-cycle, FGMRES, heat
// time integrator) that launch the gfx950 kernels.
#include "stfem_host.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <sstream>

#include <omp.h>

namespace stfem
{
  namespace
  {
    // setup phase timings on stderr when STFEM_VERBOSE is set
    struct PhaseLog
    {
      bool                                  on = std::getenv("STFEM_VERBOSE") != nullptr;
      std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
      void operator()(const char *what, long a = -1)
      {
        if (!on)
          return;
        const auto   t1 = std::chrono::steady_clock::now();
        const double s  = std::chrono::duration<double>(t1 - t0).count();
        if (a >= 0)
          std::fprintf(stderr, "[stfem] %s %ld: %.3f s\n", what, a, s);
        else
          std::fprintf(stderr, "[stfem] %s: %.3f s\n", what, s);
        t0 = t1;
      }
    };
  } // namespace

  // ========================================================== PhaseTimers
  PhaseTimers::~PhaseTimers()
  {
    for (auto &p : pending_)
      {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
      }
    for (auto e : pool_)
      (void)hipEventDestroy(e);
  }
  hipEvent_t PhaseTimers::get()
  {
    if (pool_.empty())
      {
        hipEvent_t e;
        STFEM_HIP(hipEventCreate(&e));
        return e;
      }
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t PhaseTimers::start(hipStream_t s)
  {
    hipEvent_t a = get();
    STFEM_HIP(hipEventRecord(a, s));
    return a;
  }
  void PhaseTimers::stop(int phase, hipEvent_t a, hipStream_t s)
  {
    // called from destructors: failures drop the sample instead of throwing
    hipEvent_t b = nullptr;
    try
      {
        b = get();
      }
    catch (...)
      {
        pool_.push_back(a);
        return;
      }
    if (hipEventRecord(b, s) != hipSuccess)
      {
        pool_.push_back(a);
        pool_.push_back(b);
        return;
      }
    pending_.push_back({phase, a, b});
    if (pending_.size() >= 4096)
      {
        try
          {
            resolve();
          }
        catch (...)
          {
          }
      }
  }
  void PhaseTimers::resolve()
  {
    for (auto &p : pending_)
      {
        STFEM_HIP(hipEventSynchronize(p.b));
        float t = 0;
        STFEM_HIP(hipEventElapsedTime(&t, p.a, p.b));
        ms_[p.phase] += t;
        calls_[p.phase] += 1;
        pool_.push_back(p.a);
        pool_.push_back(p.b);
      }
    pending_.clear();
  }
  void PhaseTimers::read(double *ms, long *calls)
  {
    resolve();
    for (int i = 0; i < n_phases; ++i)
      {
        if (ms)
          ms[i] = ms_[i];
        if (calls)
          calls[i] = calls_[i];
      }
  }
  void PhaseTimers::reset()
  {
    resolve();
    for (int i = 0; i < n_phases; ++i)
      {
        ms_[i]    = 0;
        calls_[i] = 0;
      }
  }
  const char *PhaseTimers::name(int phase)
  {
    static const char *n[n_phases] = {"vmult", "vanka", "gmg", "step"};
    return (phase >= 0 && phase < n_phases) ? n[phase] : "?";
  }

  // =============================================================== Context
  Context::Context(int dev)
    : device(dev)
  {
    STFEM_HIP(hipSetDevice(dev));
    STFEM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    partial.alloc(2 * size_t(reduce_slots()));
    scalars.alloc(64);
    // shape tables for every supported degree, once per device
    static std::mutex           mtx;
    static std::set<int>        done;
    std::lock_guard<std::mutex> lk(mtx);
    if (!done.count(dev))
      {
        for (int p = 1; p < MAX_N; ++p)
          {
            Shape1D sh(p);
            upload_shape_tables(sh.S.data(), sh.D.data(), sh.qw.data(), sh.qx.data(), p + 1);
          }
        done.insert(dev);
      }
  }
  Context::~Context()
  {
    if (timers)
      {
        // the reference prints its TimerOutput summary at the end
        // (printTiming); here when the problem is destroyed
        double ms[PhaseTimers::n_phases];
        long   calls[PhaseTimers::n_phases];
        try
          {
            timers->read(ms, calls);
            std::fprintf(stderr, "[stfem] phase timers (device time):\n");
            for (int i = 0; i < PhaseTimers::n_phases; ++i)
              std::fprintf(stderr, "[stfem]   %-6s %10ld calls %12.3f ms\n",
                           PhaseTimers::name(i), calls[i], ms[i]);
          }
        catch (...)
          {
          }
        timers.reset();
      }
    if (stream)
      (void)hipStreamDestroy(stream);
  }
  void Context::sync() { STFEM_HIP(hipStreamSynchronize(stream)); }

  // ========================================================== SystemMatrix
  template <typename T>
  SystemMatrix<T>::SystemMatrix(Context &ctx, const MeshLevel &m, const DMat &Alpha,
                                const DMat &Beta, bool two_pass)
    : ctx_(ctx)
    , grid_(m.g)
    , coef_mode_(m.coef_mode)
    , coef_value_(T(m.coef_value))
    , A_(Alpha)
    , B_(Beta)
    , nt_in_(Alpha.n)
    , nt_out_(Alpha.m)
  {
    STFEM_REQUIRE(Alpha.m == Beta.m && Alpha.n == Beta.n, "SystemMatrix: Alpha/Beta shapes differ");
    verts_.upload(m.X.data(), m.X.size(), ctx.stream);
    if (two_pass)
      cells_.alloc(st_cells_size(grid_, nt_out_));
    if (coef_mode_ != 0)
      {
        std::vector<T> c(m.coef.begin(), m.coef.end());
        coef_.upload(c.data(), c.size(), ctx.stream);
      }
    std::vector<T> a(A_.a.begin(), A_.a.end()), b(B_.a.begin(), B_.a.end());
    dA_.upload(a.data(), a.size(), ctx.stream);
    dB_.upload(b.data(), b.size(), ctx.stream);
    if (nt_in_ == nt_out_)
      {
        const DMat     At = A_.transposed(), Bt = B_.transposed();
        std::vector<T> at(At.a.begin(), At.a.end()), bt(Bt.a.begin(), Bt.a.end());
        dAt_.upload(at.data(), at.size(), ctx.stream);
        dBt_.upload(bt.data(), bt.size(), ctx.stream);
      }
    ctx.sync();
  }

  template <typename T>
  void SystemMatrix<T>::apply(T *dst, const T *src, const DBuf<T> &a, const DBuf<T> &b, int mode,
                              const T *rhs, T *cells) const
  {
    const PhaseScope scope(ctx_, PhaseTimers::vmult);
    StOpArgs<T> args;
    args.g          = grid_;
    args.vertices   = verts_.get();
    args.coef_mode  = coef_mode_;
    args.coef_value = coef_value_;
    args.coef       = coef_.get();
    args.alpha      = a.get();
    args.beta       = b.get();
    args.nt_in      = nt_in_;
    args.nt_out     = nt_out_;
    args.max_jb     = max_jb_ > 0 ? max_jb_ : (sizeof(T) == 8 && grid_.p >= 4 ? 2 : 4);
    args.split_below = split_below_;
    args.kernel     = op_kernel_;
    launch_st_apply<T>(args, src, dst, cells ? cells : cells_.get(), mode, rhs, ctx_.stream);
  }
  template <typename T>
  void SystemMatrix<T>::vmult_with_cells(T *dst, const T *src, T *cells) const
  {
    STFEM_REQUIRE(cells_.get() && cells, "vmult_with_cells: needs the two-pass operator");
    apply(dst, src, dA_, dB_, 0, nullptr, cells);
  }
  template <typename T>
  void SystemMatrix<T>::residual(T *dst, const T *rhs, const T *src) const
  {
    if (cells_.get())
      return apply(dst, src, dA_, dB_, 2, rhs);
    apply(dst, src, dA_, dB_, 0);
    launch_axpby<T>(m(), T(1), rhs, T(-1), dst, ctx_.stream);
  }
  template <typename T>
  void SystemMatrix<T>::vmult(T *dst, const T *src) const
  {
    apply(dst, src, dA_, dB_, 0);
  }
  template <typename T>
  void SystemMatrix<T>::vmult_cells(const T *src) const
  {
    STFEM_REQUIRE(cells_.get(), "vmult_cells: the operator has no cell buffer");
    apply(nullptr, src, dA_, dB_, 3);
  }
  template <typename T>
  void SystemMatrix<T>::set_cell_coefficient(const double *kappa)
  {
    std::vector<T> c(kappa, kappa + grid_.n_cells);
    coef_.upload(c.data(), c.size(), ctx_.stream);
    coef_mode_ = 1;
    ctx_.sync();
  }
  template <typename T>
  void SystemMatrix<T>::Tvmult(T *dst, const T *src) const
  {
    STFEM_REQUIRE(nt_in_ == nt_out_, "Tvmult needs square time matrices");
    apply(dst, src, dAt_, dBt_, 0);
  }
  template <typename T>
  void SystemMatrix<T>::vmult_add_nx1(T *dst, const T *src) const
  {
    STFEM_REQUIRE(nt_in_ == 1, "vmult_add_nx1 needs nt x 1 time matrices");
    apply(dst, src, dA_, dB_, 1);
  }
  template <typename T>
  double SystemMatrix<T>::vmult_bytes() const
  {
    const double s  = sizeof(T);
    const double nv = double(grid_.nc[0] + 1) * (grid_.nc[1] + 1) *
                      (grid_.dim == 3 ? grid_.nc[2] + 1 : 1);
    double       kb = 0;
    if (coef_mode_ == 1)
      kb = s * grid_.n_cells;
    else if (coef_mode_ == 2)
      kb = s * grid_.n_cells * std::pow(grid_.p + 1, grid_.dim);
    return s * grid_.n_nodes * (nt_in_ + nt_out_) + kb + 8.0 * grid_.dim * nv;
  }

  // ===================================================== PreconditionVanka
  namespace
  {
    // in-place Gauss-Jordan inverse with partial pivoting (row-major)
    bool gj_invert(double *A, int n, std::vector<double> &work)
    {
      work.assign(size_t(n) * n, 0.0);
      for (int i = 0; i < n; ++i)
        work[size_t(i) * n + i] = 1.0;
      for (int c = 0; c < n; ++c)
        {
          int    r = c;
          double best = std::abs(A[size_t(c) * n + c]);
          for (int k = c + 1; k < n; ++k)
            if (std::abs(A[size_t(k) * n + c]) > best)
              best = std::abs(A[size_t(k) * n + c]), r = k;
          if (best == 0.0)
            return false;
          if (r != c)
            for (int j = 0; j < n; ++j)
              {
                std::swap(A[size_t(r) * n + j], A[size_t(c) * n + j]);
                std::swap(work[size_t(r) * n + j], work[size_t(c) * n + j]);
              }
          const double d = 1.0 / A[size_t(c) * n + c];
          for (int j = 0; j < n; ++j)
            {
              A[size_t(c) * n + j] *= d;
              work[size_t(c) * n + j] *= d;
            }
          for (int k = 0; k < n; ++k)
            if (k != c)
              {
                const double f = A[size_t(k) * n + c];
                if (f == 0.0)
                  continue;
                double *Ak = A + size_t(k) * n, *Wk = work.data() + size_t(k) * n;
                const double *Ac = A + size_t(c) * n, *Wc = work.data() + size_t(c) * n;
                for (int j = 0; j < n; ++j)
                  {
                    Ak[j] -= f * Ac[j];
                    Wk[j] -= f * Wc[j];
                  }
              }
        }
      std::memcpy(A, work.data(), sizeof(double) * size_t(n) * n);
      return true;
    }

    // symmetric tridiagonal QL with implicit shifts (eigenvalues d, vectors z)
    // after Householder reduction; dense symmetric input a (row-major n x n)
    void sym_eigen(std::vector<double> &a, int n, std::vector<double> &d)
    {
      // Householder tridiagonalization (a becomes the orthogonal matrix Q)
      std::vector<double> e(n, 0.0);
      d.assign(n, 0.0);
      for (int i = n - 1; i > 0; --i)
        {
          const int l = i - 1;
          double    h = 0.0, scale = 0.0;
          if (l > 0)
            {
              for (int k = 0; k <= l; ++k)
                scale += std::abs(a[size_t(i) * n + k]);
              if (scale == 0.0)
                e[i] = a[size_t(i) * n + l];
              else
                {
                  for (int k = 0; k <= l; ++k)
                    {
                      a[size_t(i) * n + k] /= scale;
                      h += a[size_t(i) * n + k] * a[size_t(i) * n + k];
                    }
                  double f = a[size_t(i) * n + l];
                  double g = (f >= 0.0 ? -std::sqrt(h) : std::sqrt(h));
                  e[i]     = scale * g;
                  h -= f * g;
                  a[size_t(i) * n + l] = f - g;
                  f                    = 0.0;
                  for (int j = 0; j <= l; ++j)
                    {
                      a[size_t(j) * n + i] = a[size_t(i) * n + j] / h;
                      g                    = 0.0;
                      for (int k = 0; k <= j; ++k)
                        g += a[size_t(j) * n + k] * a[size_t(i) * n + k];
                      for (int k = j + 1; k <= l; ++k)
                        g += a[size_t(k) * n + j] * a[size_t(i) * n + k];
                      e[j] = g / h;
                      f += e[j] * a[size_t(i) * n + j];
                    }
                  const double hh = f / (h + h);
                  for (int j = 0; j <= l; ++j)
                    {
                      f    = a[size_t(i) * n + j];
                      e[j] = g = e[j] - hh * f;
                      for (int k = 0; k <= j; ++k)
                        a[size_t(j) * n + k] -= (f * e[k] + g * a[size_t(i) * n + k]);
                    }
                }
            }
          else
            e[i] = a[size_t(i) * n + l];
          d[i] = h;
        }
      d[0] = 0.0;
      e[0] = 0.0;
      for (int i = 0; i < n; ++i)
        {
          const int l = i - 1;
          if (d[i] != 0.0)
            for (int j = 0; j <= l; ++j)
              {
                double g = 0.0;
                for (int k = 0; k <= l; ++k)
                  g += a[size_t(i) * n + k] * a[size_t(k) * n + j];
                for (int k = 0; k <= l; ++k)
                  a[size_t(k) * n + j] -= g * a[size_t(k) * n + i];
              }
          d[i]                 = a[size_t(i) * n + i];
          a[size_t(i) * n + i] = 1.0;
          for (int j = 0; j <= l; ++j)
            a[size_t(j) * n + i] = a[size_t(i) * n + j] = 0.0;
        }
      // implicit QL on the tridiagonal (d, e), accumulating into a
      for (int i = 1; i < n; ++i)
        e[i - 1] = e[i];
      e[n - 1] = 0.0;
      for (int l = 0; l < n; ++l)
        {
          int iter = 0, m;
          do
            {
              for (m = l; m < n - 1; ++m)
                {
                  const double dd = std::abs(d[m]) + std::abs(d[m + 1]);
                  if (std::abs(e[m]) <= 1e-15 * dd)
                    break;
                }
              if (m != l)
                {
                  if (++iter > 60)
                    throw Error(STATUS_ERR_INTERNAL, "eigen: QL did not converge");
                  double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                  double r = std::hypot(g, 1.0);
                  g        = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::abs(r) : -std::abs(r)));
                  double s = 1.0, c = 1.0, pp = 0.0;
                  int    i;
                  for (i = m - 1; i >= l; --i)
                    {
                      double f = s * e[i], b = c * e[i];
                      e[i + 1] = (r = std::hypot(f, g));
                      if (r == 0.0)
                        {
                          d[i + 1] -= pp;
                          e[m] = 0.0;
                          break;
                        }
                      s        = f / r;
                      c        = g / r;
                      g        = d[i + 1] - pp;
                      r        = (d[i] - g) * s + 2.0 * c * b;
                      d[i + 1] = g + (pp = s * r);
                      g        = c * r - b;
                      for (int k = 0; k < n; ++k)
                        {
                          f                        = a[size_t(k) * n + i + 1];
                          a[size_t(k) * n + i + 1] = s * a[size_t(k) * n + i] + c * f;
                          a[size_t(k) * n + i]     = c * a[size_t(k) * n + i] - s * f;
                        }
                    }
                  if (r == 0.0 && i >= l)
                    continue;
                  d[l] -= pp;
                  e[l] = g;
                  e[m] = 0.0;
                }
            }
          while (m != l);
        }
    }
  } // namespace

  template <typename T>
  PreconditionVanka<T>::PreconditionVanka(Context &ctx, const MeshLevel &m, const Shape1D &sh,
                                          const DMat &Alpha, const DMat &Beta, Form form,
                                          int device_setup)
    : ctx_(ctx)
    , grid_(m.g)
    , nt_(Alpha.m)
    , form_(form)
  {
    if (form_ == Form::factored)
      {
        // W_k = (lambda_k Alpha + Beta)^{-1} is applied from these in the sweep
        std::vector<T> a(Alpha.a.begin(), Alpha.a.end()), b(Beta.a.begin(), Beta.a.end());
        ta_.upload(a.data(), a.size(), ctx.stream);
        tb_.upload(b.data(), b.size(), ctx.stream);
        ctx.sync();
      }
    if (form_ == Form::factored && device_setup != 0 && nt_ <= 8)
      {
        // batched generalized eigensolver: the hand-written one (Cholesky,
        // Householder and implicit QL per patch in LDS) up to 128 nodes per
        // patch, above that the reduction to a standard problem with batched
        // BLAS-3 and rocSOLVER's divide and conquer
        const int nl = (m.g.dim == 3) ? sh.n * sh.n * sh.n : sh.n * sh.n;
        int       dev = 0, lds = 0;
        STFEM_HIP(hipGetDevice(&dev));
        STFEM_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const bool own_ok = nl <= 128 && patch_eigen_lds(nl) + 64 <= size_t(lds);
        int        mth    = device_setup == 3 ? (own_ok ? 6 : 4) : device_setup;
        if (mth == 6 && !own_ok)
          mth = 4;
        setup_factored_on_device(m, sh, Alpha, Beta, mth);
        return;
      }
    std::vector<double> Kp, Mp;
    cell_patch_matrices(m, sh, Kp, Mp);
    const int    nl = (m.g.dim == 3) ? sh.n * sh.n * sh.n : sh.n * sh.n;
    const int    nt = nt_, nb = nt * nl;
    const long   nc = m.g.n_cells;
    const size_t bs = size_t(nl) * nl;
    const long   chunk = std::max<long>(1, std::min<long>(nc, (256L << 20) / (long(nb) * nb * 8)));
    if (form_ == Form::explicit_inverse)
      {
        binv_.alloc(size_t(nc) * nb * nb);
        std::vector<T> stage;
        for (long c0 = 0; c0 < nc; c0 += chunk)
          {
            const long c1 = std::min(nc, c0 + chunk);
            stage.resize(size_t(c1 - c0) * nb * nb);
            bool singular = false;
#pragma omp parallel
            {
              std::vector<double> B(size_t(nb) * nb), work;
#pragma omp for schedule(dynamic, 4)
              for (long c = c0; c < c1; ++c)
                {
                  const double *K = Kp.data() + c * bs, *M = Mp.data() + c * bs;
                  std::fill(B.begin(), B.end(), 0.0);
                  for (int i = 0; i < nt; ++i)
                    for (int j = 0; j < nt; ++j)
                      {
                        const double a = Alpha(i, j), b = Beta(i, j);
                        if (a == 0.0 && b == 0.0)
                          continue;
                        for (int k = 0; k < nl; ++k)
                          for (int l = 0; l < nl; ++l)
                            B[size_t(k + i * nl) * nb + l + j * nl] =
                              b * M[size_t(k) * nl + l] + a * K[size_t(k) * nl + l];
                      }
                  if (!gj_invert(B.data(), nb, work))
                    singular = true;
                  // column-major store: element (row, col) at col * nb + row
                  T *dst = stage.data() + size_t(c - c0) * nb * nb;
                  for (int row = 0; row < nb; ++row)
                    for (int col = 0; col < nb; ++col)
                      dst[size_t(col) * nb + row] = T(B[size_t(row) * nb + col]);
                }
            }
            if (singular)
              throw Error(STATUS_ERR_INTERNAL, "Vanka: singular patch matrix");
            STFEM_HIP(hipMemcpy(binv_.get() + size_t(c0) * nb * nb, stage.data(),
                                stage.size() * sizeof(T), hipMemcpyHostToDevice));
          }
      }
    else
      {
        STFEM_REQUIRE(nt <= 16, "Vanka factored form supports at most 16 time dofs");
        const int    ld  = vanka_s_ld(nl, sizeof(T)); // row-padded S blocks
        const size_t bsp = size_t(nl) * ld;
        S_.alloc(size_t(nc) * bsp);
        lam_.alloc(size_t(nc) * nl);
        cells_.alloc(size_t(nc) * vanka_cell_block(nl, nt, sizeof(T)));
        const long     ch = std::max<long>(1, std::min<long>(nc, 4096));
        std::vector<T> sS, sW;
        for (long c0 = 0; c0 < nc; c0 += ch)
          {
            const long c1 = std::min(nc, c0 + ch);
            sS.assign(size_t(c1 - c0) * bsp, T(0));
            sW.resize(size_t(c1 - c0) * nl);
            bool bad = false;
#pragma omp parallel
            {
              std::vector<double> Lc(bs), C(bs), lam;
#pragma omp for schedule(dynamic, 4)
              for (long c = c0; c < c1; ++c)
                {
                  const double *K = Kp.data() + c * bs, *M = Mp.data() + c * bs;
                  // Cholesky M = L L^T
                  std::fill(Lc.begin(), Lc.end(), 0.0);
                  bool ok = true;
                  for (int j = 0; j < nl && ok; ++j)
                    {
                      double s = M[size_t(j) * nl + j];
                      for (int k = 0; k < j; ++k)
                        s -= Lc[size_t(j) * nl + k] * Lc[size_t(j) * nl + k];
                      if (s <= 0)
                        {
                          ok = false;
                          break;
                        }
                      const double ljj        = std::sqrt(s);
                      Lc[size_t(j) * nl + j] = ljj;
                      for (int i = j + 1; i < nl; ++i)
                        {
                          double t = M[size_t(i) * nl + j];
                          for (int k = 0; k < j; ++k)
                            t -= Lc[size_t(i) * nl + k] * Lc[size_t(j) * nl + k];
                          Lc[size_t(i) * nl + j] = t / ljj;
                        }
                    }
                  if (!ok)
                    {
                      bad = true;
                      continue;
                    }
                  // C = L^{-1} K L^{-T}: solve L Y = K, then L C^T = Y^T
                  std::vector<double> Y(K, K + bs);
                  for (int col = 0; col < nl; ++col)
                    for (int i = 0; i < nl; ++i)
                      {
                        double s = Y[size_t(i) * nl + col];
                        for (int k = 0; k < i; ++k)
                          s -= Lc[size_t(i) * nl + k] * Y[size_t(k) * nl + col];
                        Y[size_t(i) * nl + col] = s / Lc[size_t(i) * nl + i];
                      }
                  for (int row = 0; row < nl; ++row)
                    for (int i = 0; i < nl; ++i)
                      {
                        double s = Y[size_t(row) * nl + i];
                        for (int k = 0; k < i; ++k)
                          s -= Lc[size_t(i) * nl + k] * C[size_t(row) * nl + k];
                        C[size_t(row) * nl + i] = s / Lc[size_t(i) * nl + i];
                      }
                  for (int i = 0; i < nl; ++i)
                    for (int j = i + 1; j < nl; ++j)
                      {
                        const double s = 0.5 * (C[size_t(i) * nl + j] + C[size_t(j) * nl + i]);
                        C[size_t(i) * nl + j] = C[size_t(j) * nl + i] = s;
                      }
                  sym_eigen(C, nl, lam); // C now holds the eigenvectors Q (columns)
                  // S = L^{-T} Q
                  T *Sd = sS.data() + size_t(c - c0) * bsp;
                  for (int col = 0; col < nl; ++col)
                    for (int i = nl - 1; i >= 0; --i)
                      {
                        double s = C[size_t(i) * nl + col];
                        for (int k = i + 1; k < nl; ++k)
                          s -= Lc[size_t(k) * nl + i] * double(Y[size_t(k) * nl + col]);
                        s /= Lc[size_t(i) * nl + i];
                        Y[size_t(i) * nl + col] = s;
                      }
                  for (int i = 0; i < nl; ++i)
                    for (int col = 0; col < nl; ++col)
                      Sd[size_t(i) * ld + col] = T(Y[size_t(i) * nl + col]);
                  // lambda_k; W_k = (lambda_k Alpha + Beta)^{-1} must exist
                  T *Wd = sW.data() + size_t(c - c0) * nl;
                  std::vector<double> Bk(size_t(nt) * nt), wk;
                  for (int k = 0; k < nl; ++k)
                    {
                      for (int i = 0; i < nt; ++i)
                        for (int j = 0; j < nt; ++j)
                          Bk[size_t(i) * nt + j] = lam[k] * Alpha(i, j) + Beta(i, j);
                      if (!gj_invert(Bk.data(), nt, wk))
                        bad = true;
                      Wd[k] = T(lam[k]);
                    }
                }
            }
            if (bad)
              throw Error(STATUS_ERR_INTERNAL, "Vanka factored: patch not positive definite");
            STFEM_HIP(hipMemcpy(S_.get() + size_t(c0) * bsp, sS.data(), sS.size() * sizeof(T),
                                hipMemcpyHostToDevice));
            STFEM_HIP(hipMemcpy(lam_.get() + size_t(c0) * nl, sW.data(), sW.size() * sizeof(T),
                                hipMemcpyHostToDevice));
          }
      }
  }

  // Factored Vanka on the device, in chunks of whole layers of cells (rows in
  // 2D) so that element matrices of a chunk and its neighbouring layers,
  // the patch matrices and the eigensolver workspace stay bounded:
  // element matrices -> patch submatrices K_c, M_c -> K_c Z = M_c Z Lambda
  // (rocSOLVER sygvj, Z^T M_c Z = I) -> S = Z, W_k = (lambda_k Alpha +
  // Beta)^{-1} in the level precision.
  template <typename T>
  void PreconditionVanka<T>::setup_factored_on_device(const MeshLevel &m, const Shape1D &sh,
                                                      const DMat &Alpha, const DMat &Beta,
                                                      int method)
  {
    const GridDesc &g  = m.g;
    const int       nl = (g.dim == 3) ? sh.n * sh.n * sh.n : sh.n * sh.n;
    const int       nt = nt_;
    const long      nc = g.n_cells;
    const size_t    bs = size_t(nl) * nl;
    hipStream_t     s  = ctx_.stream;
    const size_t bsp = size_t(nl) * vanka_s_ld(nl, sizeof(T)); // row-padded S block
    S_.alloc(size_t(nc) * bsp);
    lam_.alloc(size_t(nc) * nl);
    cells_.alloc(size_t(nc) * vanka_cell_block(nl, nt, sizeof(T)));
    DBuf<double> X, coef, dA, dB;
    X.upload(m.X.data(), m.X.size(), s);
    if (m.coef_mode != 0)
      coef.upload(m.coef.data(), m.coef.size(), s);
    dA.upload(Alpha.a.data(), Alpha.a.size(), s);
    dB.upload(Beta.a.data(), Beta.a.size(), s);
    DBuf<int> flag(1);
    launch_fill_zero<int>(flag.get(), 1, s);
    const long layer = (g.dim == 3) ? long(g.nc[0]) * g.nc[1] : long(g.nc[0]);
    const int  nlay  = (g.dim == 3) ? g.nc[2] : g.nc[1];
    // about 8 GB of FP64 work arrays per chunk (12 GB with the rotation log
    // of the hand-written eigensolver), at most 16384 patches
    const bool own      = method == 6;
    const long cap      = own ? patch_eigen_rotation_cap(nl) : 0;
    const long per_cell = long(bs) * 8 * 5 + cap * long(sizeof(double2));
    const long target =
      std::min<long>(16384, std::max<long>(1, ((own ? 12L : 8L) << 30) / per_cell));
    const int  lper     = int(std::max<long>(1, std::min<long>(nlay, target / layer)));
    const long emax     = std::min<long>(nlay, lper + 2) * layer, pmax = long(lper) * layer;
    DBuf<double> Me{size_t(emax) * bs}, Ke{size_t(emax) * bs}, Mp{size_t(pmax) * bs},
      Kp{size_t(pmax) * bs}, lam{size_t(pmax) * nl}, res{size_t(pmax) * nl};
    DBuf<int>     sweeps{size_t(pmax)}, info{size_t(pmax)};
    DBuf<double>  hh{own ? size_t(pmax) * nl : 0};
    DBuf<double2> rot{size_t(pmax) * cap};
    DBuf<int2>    chains{own ? size_t(pmax) * patch_eigen_chain_cap(nl) : 0};
    DBuf<int>     counts{own ? 2 * size_t(pmax) : 0};
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success)
      throw Error(STATUS_ERR_INTERNAL, "rocBLAS: handle creation failed");
    std::unique_ptr<std::remove_pointer_t<rocblas_handle>, rocblas_status (*)(rocblas_handle)>
      guard(h, rocblas_destroy_handle);
    rocblas_set_stream(h, s);
    std::vector<int> hinfo(static_cast<size_t>(pmax));
    for (int z0 = 0; z0 < nlay; z0 += lper)
      {
        const int  z1 = std::min(nlay, z0 + lper);
        const int  e_lo = std::max(0, z0 - 1), e_hi = std::min(nlay, z1 + 1);
        const long e0 = long(e_lo) * layer, ne = long(e_hi - e_lo) * layer;
        const long p0 = long(z0) * layer, np = long(z1 - z0) * layer;
        launch_element_matrices(g, X.get(), m.coef_mode, m.coef_value, coef.get(), sh.S.data(),
                                sh.Sd.data(), sh.qx.data(), sh.qw.data(), e0, ne, Me.get(),
                                Ke.get(), s);
        launch_patch_matrices(g, p0, np, e0, Me.get(), Ke.get(), Mp.get(), Kp.get(), s);
        // K_c z = lambda M_c z with z^T M_c z = 1: eigenvectors overwrite Kp
        rocblas_status st = rocblas_status_success;
        int            mth = method;
        if (own)
          {
            launch_patch_eigen(nl, np, Mp.get(), Kp.get(), lam.get(), res.get(), hh.get(),
                               rot.get(), chains.get(), counts.get(), info.get(), s);
            STFEM_HIP(hipMemcpyAsync(hinfo.data(), info.get(), size_t(np) * sizeof(int),
                                     hipMemcpyDeviceToHost, s));
            ctx_.sync();
            bool redo = false;
            for (long k = 0; k < np; ++k)
              {
                if (hinfo[k] == 1)
                  throw Error(STATUS_ERR_INTERNAL,
                              "Vanka factored: patch mass not positive definite");
                redo = redo || hinfo[k] != 0;
              }
            if (redo)
              {
                // an eigenvalue iteration that did not settle (or a full
                // rotation log): this chunk again by divide and conquer
                launch_patch_matrices(g, p0, np, e0, Me.get(), Ke.get(), Mp.get(), Kp.get(), s);
                mth = 4;
              }
          }
        if (own && mth == 6)
          {
            // eigenvectors in Kp, eigenvalues in lam
          }
        else if (mth >= 4)
          {
            // reduction to a standard problem with batched BLAS-3:
            // M = L L^T, C = L^{-1} K L^{-T}, C Q = Q Lambda, Z = L^{-T} Q
            // (C Q = Q Lambda by divide and conquer, method 4, or Jacobi, 5)
            const double   one = 1.0;
            const rocblas_stride sb = rocblas_stride(bs);
            const rocblas_int    nb = rocblas_int(np);
            st = rocsolver_dpotrf_strided_batched(h, rocblas_fill_lower, nl, Mp.get(), nl, sb,
                                                  info.get(), nb);
            if (st == rocblas_status_success)
              st = rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_lower,
                                                 rocblas_operation_none, rocblas_diagonal_non_unit,
                                                 nl, nl, &one, Mp.get(), nl, sb, Kp.get(), nl, sb,
                                                 nb);
            if (st == rocblas_status_success)
              st = rocblas_dtrsm_strided_batched(h, rocblas_side_right, rocblas_fill_lower,
                                                 rocblas_operation_transpose,
                                                 rocblas_diagonal_non_unit, nl, nl, &one,
                                                 Mp.get(), nl, sb, Kp.get(), nl, sb, nb);
            if (st == rocblas_status_success)
              {
                STFEM_HIP(hipMemcpyAsync(hinfo.data(), info.get(), size_t(np) * sizeof(int),
                                         hipMemcpyDeviceToHost, s));
                ctx_.sync();
                for (long k = 0; k < np; ++k)
                  if (hinfo[k] != 0)
                    throw Error(STATUS_ERR_INTERNAL, "Vanka factored: patch mass not positive definite");
                st = mth == 4 ?
                       rocsolver_dsyevd_strided_batched(h, rocblas_evect_original,
                                                        rocblas_fill_lower, nl, Kp.get(), nl, sb,
                                                        lam.get(), nl, res.get(), nl, info.get(),
                                                        nb) :
                       rocsolver_dsyevj_strided_batched(h, rocblas_esort_ascending,
                                                        rocblas_evect_original,
                                                        rocblas_fill_lower, nl, Kp.get(), nl, sb,
                                                        0.0, res.get(), 100, sweeps.get(),
                                                        lam.get(), nl, info.get(), nb);
              }
            if (st == rocblas_status_success)
              st = rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_lower,
                                                 rocblas_operation_transpose,
                                                 rocblas_diagonal_non_unit, nl, nl, &one,
                                                 Mp.get(), nl, sb, Kp.get(), nl, sb, nb);
          }
        else if (method == 2)
          st = rocsolver_dsygvd_strided_batched(h, rocblas_eform_ax, rocblas_evect_original,
                                                rocblas_fill_upper, nl, Kp.get(), nl,
                                                rocblas_stride(bs), Mp.get(), nl,
                                                rocblas_stride(bs), lam.get(), nl, res.get(), nl,
                                                info.get(), rocblas_int(np));
        else
          st = rocsolver_dsygvj_strided_batched(h, rocblas_eform_ax, rocblas_evect_original,
                                                rocblas_fill_upper, nl, Kp.get(), nl,
                                                rocblas_stride(bs), Mp.get(), nl,
                                                rocblas_stride(bs), 0.0, res.get(), 100,
                                                sweeps.get(), lam.get(), nl, info.get(),
                                                rocblas_int(np));
        if (st != rocblas_status_success)
          throw Error(STATUS_ERR_INTERNAL,
                      std::string("Vanka eigensolver: ") + rocblas_status_to_string(st));
        launch_pack_factors<T>(np, nl, nt, Kp.get(), lam.get(), dA.get(), dB.get(),
                               S_.get() + size_t(p0) * bsp, lam_.get() + size_t(p0) * nl,
                               flag.get(), s);
        STFEM_HIP(hipMemcpyAsync(hinfo.data(), info.get(), size_t(np) * sizeof(int),
                                 hipMemcpyDeviceToHost, s));
        ctx_.sync();
        for (long k = 0; k < np; ++k)
          if (hinfo[k] != 0)
            throw Error(STATUS_ERR_INTERNAL, "Vanka factored: patch not positive definite");
      }
    int hf = 0;
    STFEM_HIP(hipMemcpyAsync(&hf, flag.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    ctx_.sync();
    if (hf)
      throw Error(STATUS_ERR_INTERNAL, "Vanka factored: singular time block");
  }

  template <typename T>
  void PreconditionVanka<T>::vmult(T *dst, const T *src) const
  {
    if (form_ == Form::factored)
      {
        const PhaseScope scope(ctx_, PhaseTimers::vanka);
        return launch_vanka_factored<T>(grid_, nt_, S_.get(), lam_.get(), ta_.get(), tb_.get(), src,
                                      nullptr, dst, T(1), cells_.get(), 0, variant_, ctx_.stream);
      }
    launch_fill_zero<T>(dst, size_t(grid_.n_nodes) * nt_, ctx_.stream);
    relax(dst, src, nullptr, T(1));
  }
  template <typename T>
  void PreconditionVanka<T>::relax(T *u, const T *rhs, const T *ax, T omega, bool assign) const
  {
    const PhaseScope scope(ctx_, PhaseTimers::vanka);
    if (form_ == Form::explicit_inverse)
      {
        if (assign)
          launch_fill_zero<T>(u, size_t(grid_.n_nodes) * nt_, ctx_.stream);
        launch_vanka_explicit<T>(grid_, nt_, binv_.get(), rhs, ax, u, omega, ctx_.stream);
      }
    else
      launch_vanka_factored<T>(grid_, nt_, S_.get(), lam_.get(), ta_.get(), tb_.get(), rhs, ax, u,
                               omega, cells_.get(), assign ? 0 : 1, variant_, ctx_.stream);
  }
  template <typename T>
  bool PreconditionVanka<T>::fused_residual_ok() const
  {
    // the register-tiled kernel has the fused form (chosen automatically
    // from 64-node patches on); the others keep the separate gather
    const int nl = grid_.dim == 3 ? (grid_.p + 1) * (grid_.p + 1) * (grid_.p + 1) :
                                    (grid_.p + 1) * (grid_.p + 1);
    return form_ == Form::factored && cells_.get() &&
           (variant_ == 3 || (variant_ == 0 && nl >= 64)) &&
           vanka_fused_residual_supported<T>(grid_, nt_);
  }
  template <typename T>
  void PreconditionVanka<T>::cell_pass(const T *src) const
  {
    STFEM_REQUIRE(form_ == Form::factored && cells_.get(),
                  "Vanka cell pass: needs the factored form with the two-pass buffer");
    launch_vanka_factored<T>(grid_, nt_, S_.get(), lam_.get(), ta_.get(), tb_.get(), src, nullptr,
                             nullptr, T(1), cells_.get(), 3, variant_, ctx_.stream);
  }
  template <typename T>
  void PreconditionVanka<T>::relax_opc(T *u, const T *rhs, const T *opc, T omega) const
  {
    const PhaseScope scope(ctx_, PhaseTimers::vanka);
    launch_vanka_factored<T>(grid_, nt_, S_.get(), lam_.get(), ta_.get(), tb_.get(), rhs, nullptr,
                             u, omega, cells_.get(), 1, variant_, ctx_.stream, opc);
  }
  template <typename T>
  double PreconditionVanka<T>::sweep_bytes() const
  {
    const double s  = sizeof(T);
    const double nl = std::pow(grid_.p + 1, grid_.dim);
    const double nst = double(grid_.n_nodes) * nt_;
    if (form_ == Form::explicit_inverse)
      return s * grid_.n_cells * (nt_ * nl) * (nt_ * nl) + 3 * s * nst;
    return s * grid_.n_cells * (nl * nl + nl) + 3 * s * nst;
  }

  // ===================================================== time transfer
  template <typename T>
  MGTwoLevelTransferST<T>::MGTwoLevelTransferST(Context &ctx, TimeStepType type, int r, int nts,
                                                TimeMGType kind, long n_nodes, bool overwrite,
                                                bool restrict_is_transpose)
    : ctx_(ctx)
    , n_nodes_(n_nodes)
    , overwrite_(overwrite)
  {
    P_ = (kind == TimeMGType::k) ? get_time_projection_matrix(type, r - 1, r, nts) :
                                   get_time_prolongation_matrix(type, r, nts);
    // gmg.h:101-113: R = P^T, or the L2 restriction / projection onto r - 1
    const DMat R = restrict_is_transpose ? P_.transposed() :
                   (kind == TimeMGType::k) ? get_time_projection_matrix(type, r, r - 1, nts) :
                                             get_time_restriction_matrix(type, r, nts);
    STFEM_REQUIRE(R.m == P_.n && R.n == P_.m, "time transfer: restriction has the wrong shape");
    std::vector<T> p(P_.a.begin(), P_.a.end()), rr(R.a.begin(), R.a.end());
    dP_.upload(p.data(), p.size(), ctx.stream);
    dR_.upload(rr.data(), rr.size(), ctx.stream);
    ctx.sync();
  }
  template <typename T>
  void MGTwoLevelTransferST<T>::prolongate_and_add(T *dst, const T *src) const
  {
    launch_time_transfer<T>(n_nodes_, dP_.get(), P_.m, P_.n, src, dst, !overwrite_, ctx_.stream);
  }
  template <typename T>
  void MGTwoLevelTransferST<T>::restrict_and_add(T *dst, const T *src) const
  {
    launch_time_transfer<T>(n_nodes_, dR_.get(), P_.n, P_.m, src, dst, !overwrite_, ctx_.stream);
  }

  // ===================================================== space transfer
  template <typename T>
  MGTwoLevelBlockTransfer<T>::MGTwoLevelBlockTransfer(Context &ctx, const GridDesc &fine,
                                                      const GridDesc &coarse, const Shape1D &sh,
                                                      int nt)
    : ctx_(ctx)
    , fine_(fine)
    , coarse_(coarse)
    , nt_(nt)
  {
    const int p = sh.p;
    for (int d = 0; d < fine.dim; ++d)
      {
        STFEM_REQUIRE(fine.nc[d] == 2 * coarse.nc[d], "space transfer: levels not nested");
        const int nf = fine.nn[d], ncn = coarse.nn[d];
        const bool lo_d = (fine.dirichlet >> (2 * d)) & 1, hi_d = (fine.dirichlet >> (2 * d + 1)) & 1;
        // 1D prolongation rows
        std::vector<std::vector<std::pair<int, double>>> rows(nf);
        for (int f = 0; f < nf; ++f)
          {
            if ((f == 0 && lo_d) || (f == nf - 1 && hi_d))
              continue;
            int fc = f / p;
            int a  = f - fc * p;
            if (fc >= fine.nc[d])
              fc = fine.nc[d] - 1, a = p;
            const int    cc = fc / 2, child = fc % 2;
            const double xi = 0.5 * (sh.gll[a] + child);
            for (int b = 0; b <= p; ++b)
              {
                const double w   = sh.L.value(b, xi);
                const int    col = cc * p + b;
                if (std::abs(w) < 1e-14)
                  continue;
                if ((col == 0 && lo_d) || (col == ncn - 1 && hi_d))
                  continue;
                rows[f].push_back({col, w});
              }
          }
        std::vector<std::vector<std::pair<int, double>>> cols(ncn);
        for (int f = 0; f < nf; ++f)
          for (auto &e : rows[f])
            cols[e.first].push_back({f, e.second});
        auto pack = [&](Ell &E, const std::vector<std::vector<std::pair<int, double>>> &R) {
          E.rows = int(R.size());
          E.K    = 1;
          for (auto &r : R)
            E.K = std::max<int>(E.K, int(r.size()));
          std::vector<int> cnt(E.rows), col(size_t(E.rows) * E.K, 0);
          std::vector<T>   w(size_t(E.rows) * E.K, T(0));
          for (int i = 0; i < E.rows; ++i)
            {
              cnt[i] = int(R[i].size());
              for (size_t k = 0; k < R[i].size(); ++k)
                {
                  col[size_t(i) * E.K + k] = R[i][k].first;
                  w[size_t(i) * E.K + k]   = T(R[i][k].second);
                }
            }
          E.cnt.upload(cnt.data(), cnt.size(), ctx.stream);
          E.col.upload(col.data(), col.size(), ctx.stream);
          E.w.upload(w.data(), w.size(), ctx.stream);
        };
        pack(P_[d], rows);
        pack(R_[d], cols);
      }
    // temporaries: largest intermediate of the sweeps
    const long f0 = fine.nn[0], f1 = fine.nn[1], f2 = fine.nn[2];
    const long c0 = coarse.nn[0], c1 = coarse.nn[1], c2 = coarse.nn[2];
    const long big = std::max(std::max(f0 * c1 * c2, f0 * f1 * c2), std::max(c0 * f1 * f2, c0 * c1 * f2));
    tmp1_.alloc(size_t(big) * nt);
    tmp2_.alloc(size_t(big) * nt);
    ctx.sync();
  }

  template <typename T>
  void MGTwoLevelBlockTransfer<T>::prolongate_and_add(T *dst, const T *src) const
  {
    int        n[3] = {coarse_.nn[0], coarse_.nn[1], coarse_.nn[2]};
    const int  D    = fine_.dim;
    const T   *in   = src;
    T         *bufs[2] = {const_cast<T *>(tmp1_.get()), const_cast<T *>(tmp2_.get())};
    for (int d = 0; d < D; ++d)
      {
        const bool last = d == D - 1;
        T         *out  = last ? dst : bufs[d % 2];
        launch_sweep<T>(in, out, nt_, d, n, fine_.nn[d], P_[d].cnt.get(), P_[d].col.get(),
                        P_[d].w.get(), P_[d].K, last, ctx_.stream);
        n[d] = fine_.nn[d];
        in   = out;
      }
  }
  template <typename T>
  void MGTwoLevelBlockTransfer<T>::restrict_and_add(T *dst, const T *src) const
  {
    int        n[3] = {fine_.nn[0], fine_.nn[1], fine_.nn[2]};
    const int  D    = fine_.dim;
    const T   *in   = src;
    T         *bufs[2] = {const_cast<T *>(tmp1_.get()), const_cast<T *>(tmp2_.get())};
    for (int d = 0; d < D; ++d)
      {
        const bool last = d == D - 1;
        T         *out  = last ? dst : bufs[d % 2];
        launch_sweep<T>(in, out, nt_, d, n, coarse_.nn[d], R_[d].cnt.get(), R_[d].col.get(),
                        R_[d].w.get(), R_[d].K, last, ctx_.stream);
        n[d] = coarse_.nn[d];
        in   = out;
      }
  }

  // =================================================================== GMG
  template <typename T>
  GMG<T>::GMG(Context &ctx, std::vector<GMGLevel<T>> &&levels,
              std::vector<TwoLevelTransfer<T>> &&transfers, const GMGOptions &opt,
              const std::vector<double> *omega)
    : ctx_(ctx)
    , lv_(std::move(levels))
    , tr_(std::move(transfers))
    , opt_(opt)
  {
    const int nl = int(lv_.size());
    sol_.resize(nl);
    def_.resize(nl);
    t1_.resize(nl);
    t2_.resize(nl);
    for (int l = 0; l < nl; ++l)
      {
        const size_t sz = size_t(lv_[l].A->m());
        sol_[l].alloc(sz);
        def_[l].alloc(sz);
        t1_[l].alloc(sz);
        t2_[l].alloc(sz);
      }
    STFEM_REQUIRE(!omega || int(omega->size()) == nl, "GMG: one relaxation parameter per level");
    for (int l = 0; l < nl; ++l)
      lv_[l].omega = omega ? (*omega)[l] : opt_.estimate_relaxation ? estimate_omega(l) : 1.0;
    setup_dense_coarse();
    // timed phases record events around every level call: no graph then
    use_graph_ = opt_.use_graph == 1 && !ctx_.timers;
    if (opt_.use_graph == -1 && !ctx_.timers)
      {
        // rocprofv3 loads its tool library through ROCP_TOOL_LIBRARIES; with
        // a tool present the V-cycle is launched eagerly (same kernels)
        const char *tool = std::getenv("ROCP_TOOL_LIBRARIES");
        use_graph_       = !(tool && *tool);
        if (!use_graph_)
          {
            static std::once_flag once;
            std::call_once(once, [] {
              std::fprintf(stderr, "[stfem] profiler tool loaded: the V-cycle is launched "
                                   "kernel by kernel instead of as a HIP graph\n");
            });
          }
      }
  }

  namespace
  {
    // column-major FP64 products on the device (setup of the dense coarse
    // operator only)
    struct DeviceBlas
    {
      rocblas_handle h = nullptr;
      explicit DeviceBlas(hipStream_t s)
      {
        check(rocblas_create_handle(&h));
        check(rocblas_set_stream(h, s));
      }
      ~DeviceBlas()
      {
        if (h)
          (void)rocblas_destroy_handle(h);
      }
      static void check(rocblas_status st)
      {
        if (st != rocblas_status_success)
          throw Error(STATUS_ERR_INTERNAL, std::string("rocBLAS: ") + rocblas_status_to_string(st));
      }
      // C (m x n) = alpha A (m x k) B (k x n) + beta C
      void gemm(long m, long n, long k, double alpha, const double *A, const double *B, double beta,
                double *C)
      {
        check(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, rocblas_int(m),
                            rocblas_int(n), rocblas_int(k), &alpha, A, rocblas_int(m), B,
                            rocblas_int(k), &beta, C, rocblas_int(m)));
      }
      void gemm(long m, long n, long k, float alpha, const float *A, const float *B, float beta,
                float *C)
      {
        check(rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, rocblas_int(m),
                            rocblas_int(n), rocblas_int(k), &alpha, A, rocblas_int(m), B,
                            rocblas_int(k), &beta, C, rocblas_int(m)));
      }
    };
  } // namespace

  // The V-cycle on levels 0..lc is linear in def[lc]: sol[lc] = M_lc def[lc]
  // with (level_v_step)
  //   M_0 = S_0,
  //   M_l = G_l^s (S_l + P_l M_{l-1} R_l (I - A_l S_l)) + S_l
  //         (S_l dropped from the bracket when the prolongation overwrites),
  // where one relaxation sweep is u <- G u + H def, G = I - w V A, H = w V,
  // s = 2^(L-l) * smoothing_steps sweeps, and S_l = sum_{k<s} G^k H is the
  // map of s sweeps from zero (one sweep with the exact-level weight on
  // one-cell levels).  A, V, P, R are the level kernels' matrices (probed
  // with unit vectors, below), so M_lc is the V-cycle of the launched
  // kernels up to rounding; the products run in the level precision (FP32
  // levels: rocBLAS SGEMM, FP64 levels: DGEMM), in which M_lc is stored.
  template <typename T>
  void GMG<T>::setup_dense_coarse()
  {
    dense_level_ = -1;
    const int L  = int(lv_.size()) - 1;
    if (opt_.dense_coarse_dofs <= 0 || L < 1)
      return;
    int lc = -1;
    for (int l = 0; l < L; ++l)
      {
        if (lv_[l].A->m() > opt_.dense_coarse_dofs)
          break;
        lc = l;
      }
    if (lc < 0)
      return;
    hipStream_t s = ctx_.stream;
    DeviceBlas  blas(s);
    // dense n_out x n_in matrix of a local linear map (structured grids)
    // given by its launcher, by probing: the inputs of
    // one colour (lattice position = c mod stride per direction, one time
    // component) have disjoint output supports, so one application gives
    // all their columns; stride * sn > 2 * sd * radius, where an input at
    // lattice position g reaches the outputs o with |sn g - sd o| <= sd radius
    DBuf<int> bad(1);
    STFEM_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int), s));
    auto dense_probe = [&](const GridDesc &gin, int nt_in, const GridDesc &gout, int nt_out,
                           int stride, int sn, int sd, int radius, auto &&apply) {
      const long   n_in = gin.n_nodes * nt_in, n_out = gout.n_nodes * nt_out;
      DBuf<T>    D(size_t(n_out) * n_in);
      DBuf<T>    x(n_in), y(n_out);
      launch_fill_zero<T>(D.get(), D.size(), s);
      int       c[3] = {0, 0, 0};
      const int cz   = gin.dim > 2 ? std::min(stride, gin.nn[2]) : 1;
      const int cy   = std::min(stride, gin.nn[1]), cx = std::min(stride, gin.nn[0]);
      for (c[2] = 0; c[2] < cz; ++c[2])
        for (c[1] = 0; c[1] < cy; ++c[1])
          for (c[0] = 0; c[0] < cx; ++c[0])
            for (int t = 0; t < nt_in; ++t)
              {
                launch_probe_vector<T>(gin, nt_in, c, t, stride, x.get(), s);
                apply(y.get(), x.get());
                launch_probe_store<T, T>(gin, nt_in, gout, nt_out, c, t, stride, sn, sd, radius,
                                         y.get(), D.get(), n_out, bad.get(), s);
              }
      return D;
    };
    auto copy_of = [&](const DBuf<T> &a) {
      DBuf<T> b(a.size());
      STFEM_HIP(hipMemcpyAsync(b.get(), a.get(), a.size() * sizeof(T),
                               hipMemcpyDeviceToDevice, s));
      return b;
    };
    auto identity = [&](long n) {
      DBuf<T> I(size_t(n) * n);
      launch_fill_zero<T>(I.get(), I.size(), s);
      launch_add_diagonal<T>(n, T(1), I.get(), s);
      return I;
    };
    DBuf<T> Mc;
    long    nc = 0;
    for (int l = 0; l <= lc; ++l)
      {
        const long n = lv_[l].A->m();
        const auto &Lv = lv_[l];
        // A and V couple the nodes of the cells around a node: |o - g| <= p
        const GridDesc &gl = Lv.A->grid();
        const int       ntl = Lv.A->n_blocks(), pl = gl.p;
        DBuf<T> A = dense_probe(gl, ntl, gl, ntl, 2 * pl + 1, 1, 1, pl,
                                [&](T *y, const T *x) { Lv.A->vmult(y, x); });
        DBuf<T> V = dense_probe(gl, ntl, gl, ntl, 2 * pl + 1, 1, 1, pl,
                                [&](T *y, const T *x) { Lv.V->vmult(y, x); });
        const int sweeps = steps(l) * opt_.smoothing_steps;
        // (Gs, S): s sweeps; G = I - w V A is the one-sweep iteration matrix
        auto sweep_pair = [&](double w) {
          DBuf<T> G = identity(n);
          blas.gemm(n, n, n, T(-w), V.get(), A.get(), T(1), G.get());
          DBuf<T> H = copy_of(V);
          launch_scale<T>(long(H.size()), T(w), H.get(), s);
          return std::make_pair(std::move(G), std::move(H));
        };
        DBuf<T> Gs, S;
        if (exact_level(l))
          {
            auto gh = sweep_pair(double(exact_omega(l, sweeps)));
            Gs      = std::move(gh.first);
            S       = std::move(gh.second);
          }
        else
          {
            auto gh = sweep_pair(double(T(lv_[l].omega)));
            DBuf<T> Gb = std::move(gh.first), Sb = std::move(gh.second);
            bool    have = false;
            for (int k = sweeps; k > 0; k >>= 1)
              {
                if (k & 1)
                  {
                    if (!have)
                      {
                        Gs   = copy_of(Gb);
                        S    = copy_of(Sb);
                        have = true;
                      }
                    else
                      {
                        // (Gs, S) followed by the base block (Gb, Sb)
                        DBuf<T> S2 = copy_of(Sb);
                        blas.gemm(n, n, n, T(1), Gb.get(), S.get(), T(1), S2.get());
                        DBuf<T> G2(size_t(n) * n);
                        blas.gemm(n, n, n, T(1), Gb.get(), Gs.get(), T(0), G2.get());
                        S  = std::move(S2);
                        Gs = std::move(G2);
                      }
                  }
                if (k > 1)
                  {
                    DBuf<T> S2 = copy_of(Sb);
                    blas.gemm(n, n, n, T(1), Gb.get(), Sb.get(), T(1), S2.get());
                    DBuf<T> G2(size_t(n) * n);
                    blas.gemm(n, n, n, T(1), Gb.get(), Gb.get(), T(0), G2.get());
                    Sb = std::move(S2);
                    Gb = std::move(G2);
                  }
              }
          }
        if (l == 0)
          {
            Mc = std::move(S);
            nc = n;
            continue;
          }
        const TwoLevelTransfer<T> &tr = tr_[l];
        const GridDesc &gc  = lv_[l - 1].A->grid();
        const int       ntc = lv_[l - 1].A->n_blocks();
        auto prolong = [&](T *y, const T *x) {
          launch_fill_zero<T>(y, size_t(n), s);
          tr.prolongate_and_add(y, x);
        };
        DBuf<T> P, R;
        if (tr.time)
          {
            // time transfers act node by node
            P = dense_probe(gc, ntc, gl, ntl, 1, 1, 1, 0, prolong);
            R = dense_probe(gl, ntl, gc, ntc, 1, 1, 1, 0, [&](T *y, const T *x) {
              launch_fill_zero<T>(y, size_t(nc), s);
              tr.restrict_and_add(y, x);
            });
          }
        else
          {
            // h: a coarse node reaches the fine nodes of its coarse cells,
            // |o - 2 g| <= 2p; the restriction is the transpose (the same 1D
            // weights, gmg.h:371-380)
            P = dense_probe(gc, ntc, gl, ntl, 2 * pl + 1, 2, 1, 2 * pl, prolong);
            R = DBuf<T>(size_t(nc) * n);
            launch_transpose<T>(n, nc, P.get(), R.get(), s);
          }
        DBuf<T> X = identity(n); // I - A S
        blas.gemm(n, n, n, T(-1), A.get(), S.get(), T(1), X.get());
        DBuf<T> Y(size_t(nc) * n); // R X
        blas.gemm(nc, n, n, T(1), R.get(), X.get(), T(0), Y.get());
        DBuf<T> Z(size_t(nc) * n); // M_{l-1} Y
        blas.gemm(nc, n, nc, T(1), Mc.get(), Y.get(), T(0), Z.get());
        DBuf<T> U = tr.prolongation_overwrites() ? DBuf<T>(size_t(n) * n) : copy_of(S);
        blas.gemm(n, n, nc, T(1), P.get(), Z.get(), tr.prolongation_overwrites() ? T(0) : T(1),
                  U.get());
        DBuf<T> M = copy_of(S); // Gs U + S
        blas.gemm(n, n, n, T(1), Gs.get(), U.get(), T(1), M.get());
        Mc = std::move(M);
        nc = n;
      }
    int hbad = 0;
    STFEM_HIP(hipMemcpyAsync(&hbad, bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    ctx_.sync();
    STFEM_REQUIRE(hbad == 0, "dense coarse operator: probed columns overlap");
    const long vw = long(16 / sizeof(T));
    dense_ld_     = (nc + vw - 1) / vw * vw;
    dense_rows_.alloc(size_t(nc) * dense_ld_);
    launch_pack_rows<T, T>(nc, Mc.get(), dense_rows_.get(), dense_ld_, s);
    ctx_.sync();
    dense_level_ = lc;
  }

  template <typename T>
  int GMG<T>::steps(int l) const
  {
    const int L = int(lv_.size()) - 1;
    return opt_.variable ? (1 << (L - l)) : 1;
  }

  // PreconditionChebyshev::estimate_eigenvalues (power iteration on
  // Vanka^{-1} A, constant start vector), omega = 2 / (alpha + lambda_max),
  // alpha = 0.9 lambda_max (gmg.h:792-844).
  template <typename T>
  double GMG<T>::estimate_omega(int l) const
  {
    if (l == 0)
      return 1.0;
    const auto  &L  = lv_[l];
    const long   sz = L.A->m();
    T           *ev = sol_[l].get(), *v1 = t1_[l].get(), *v2 = t2_[l].get();
    // start vector: ones, zero at the first unknown (node 0, time 0) and at
    // the Dirichlet nodes
    std::vector<T> ones(sz, T(1));
    ones[0] = T(0);
    STFEM_HIP(hipMemcpyAsync(ev, ones.data(), sz * sizeof(T), hipMemcpyHostToDevice, ctx_.stream));
    launch_zero_dirichlet<T>(L.A->grid(), L.A->n_blocks(), ev, ctx_.stream);
    double *dsc = ctx_.scalars.get();
    double  h[2];
    launch_dot<T>(sz, ev, ev, ctx_.partial.get(), dsc, 1, ctx_.stream);
    STFEM_HIP(hipMemcpyAsync(h, dsc, sizeof(double), hipMemcpyDeviceToHost, ctx_.stream));
    ctx_.sync();
    if (!(h[0] > 0.0))
      return 1.0; // no free unknowns on this level (all nodes on Dirichlet faces)
    launch_axpby<T>(sz, T(1.0 / h[0]), ev, T(0), ev, ctx_.stream);
    double est = 0;
    for (int it = 0; it < opt_.eig_iterations; ++it)
      {
        L.A->vmult(v2, ev);
        L.V->vmult(v1, v2);
        launch_dot<T>(sz, ev, v1, ctx_.partial.get(), dsc, 0, ctx_.stream);
        launch_dot<T>(sz, v1, v1, ctx_.partial.get(), dsc + 1, 1, ctx_.stream);
        STFEM_HIP(hipMemcpyAsync(h, dsc, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx_.stream));
        ctx_.sync();
        est = h[0];
        launch_axpby<T>(sz, T(1.0 / h[1]), v1, T(0), ev, ctx_.stream);
      }
    if (!std::isfinite(est) || est == 0.0)
      return 1.0;
    const double lmax = std::abs(est) * opt_.eig_safety;
    return 2.0 / (0.9 * lmax + lmax);
  }

  template <typename T>
  GMG<T>::~GMG()
  {
    if (exec_)
      (void)hipGraphExecDestroy(exec_);
    if (graph_)
      (void)hipGraphDestroy(graph_);
  }

  // One PreconditionRelaxation sweep with the Vanka preconditioner:
  // first sweep from zero u = omega V rhs, later sweeps
  // u += omega V (rhs - A u).  Two kernels per sweep (A u, fused Vanka
  // update) plus a zero fill.
  template <typename T>
  void GMG<T>::relax(int l, T *u, const T *rhs, bool first) const
  {
    relax_omega(l, u, rhs, first, T(lv_[l].omega));
  }
  template <typename T>
  void GMG<T>::relax_omega(int l, T *u, const T *rhs, bool first, T omega) const
  {
    const auto &L  = lv_[l];
    const long  sz = L.A->m();
    (void)sz;
    if (first)
      {
        L.V->relax(u, rhs, nullptr, omega, true);
        return;
      }
    if (L.V->n_cells() <= opt_.fuse_residual && L.A->cell_buffer() && L.V->fused_residual_ok())
      {
        L.A->vmult_cells(u);
        L.V->relax_opc(u, rhs, L.A->cell_buffer(), omega);
        return;
      }
    T *t1 = t1_[l].get();
    L.A->vmult(t1, u);
    L.V->relax(u, rhs, t1, omega);
  }
  template <typename T>
  bool GMG<T>::exact_level(int l) const
  {
    return opt_.exact_single_cell && lv_[l].V->n_cells() == 1;
  }
  // total sweeps s: on a one-cell level V = A^{-1}, so the error contracts
  // by (1 - omega) per sweep and s sweeps equal one with 1 - (1 - omega)^s
  template <typename T>
  T GMG<T>::exact_omega(int l, int sweeps) const
  {
    return T(1.0 - std::pow(1.0 - lv_[l].omega, double(sweeps)));
  }
  template <typename T>
  void GMG<T>::smoother_apply(int l, T *u, const T *rhs) const
  {
    const int s2 = steps(l);
    if (exact_level(l))
      return relax_omega(l, u, rhs, true, exact_omega(l, s2 * opt_.smoothing_steps));
    for (int i = 0; i < opt_.smoothing_steps; ++i)
      relax(l, u, rhs, i == 0);
    for (int k = 1; k < s2; ++k)
      for (int i = 1; i <= opt_.smoothing_steps; ++i)
        relax(l, u, rhs, false);
  }
  template <typename T>
  void GMG<T>::smoother_smooth(int l, T *u, const T *rhs) const
  {
    const int s2 = steps(l);
    if (exact_level(l))
      return relax_omega(l, u, rhs, false, exact_omega(l, s2 * opt_.smoothing_steps));
    for (int k = 0; k < s2; ++k)
      for (int i = 1; i <= opt_.smoothing_steps; ++i)
        relax(l, u, rhs, false);
  }
  template <typename T>
  void GMG<T>::level_v_step(int l) const
  {
    if (l == dense_level_)
      {
        launch_gemv_rows<T>(lv_[l].A->m(), dense_ld_, dense_rows_.get(), def_[l].get(),
                            sol_[l].get(), ctx_.stream);
        return;
      }
    if (l == 0)
      {
        smoother_apply(0, sol_[0].get(), def_[0].get());
        return;
      }
    smoother_apply(l, sol_[l].get(), def_[l].get());
    T *t = t1_[l].get();
    lv_[l].A->residual(t, def_[l].get(), sol_[l].get());
    launch_fill_zero<T>(def_[l - 1].get(), def_[l - 1].size(), ctx_.stream);
    tr_[l].restrict_and_add(def_[l - 1].get(), t);
    level_v_step(l - 1);
    tr_[l].prolongate_and_add(sol_[l].get(), sol_[l - 1].get());
    smoother_smooth(l, sol_[l].get(), def_[l].get());
  }
  template <typename T>
  void GMG<T>::vcycle() const
  {
    level_v_step(int(lv_.size()) - 1);
  }
  template <typename T>
  void GMG<T>::vmult(double *dst, const double *src) const
  {
    const PhaseScope scope(ctx_, PhaseTimers::gmg);
    const int  L  = int(lv_.size()) - 1;
    const long sz = lv_[L].A->m();
    if constexpr (std::is_same_v<T, double>)
      STFEM_HIP(hipMemcpyAsync(def_[L].get(), src, sz * sizeof(double), hipMemcpyDeviceToDevice,
                               ctx_.stream));
    else
      launch_convert<T, double>(sz, src, def_[L].get(), ctx_.stream);
    if (use_graph_)
      {
        // the V-cycle launches a fixed kernel sequence on fixed buffers:
        // record it once, replay it afterwards
        if (!exec_)
          {
            STFEM_HIP(hipStreamBeginCapture(ctx_.stream, hipStreamCaptureModeThreadLocal));
            vcycle();
            STFEM_HIP(hipStreamEndCapture(ctx_.stream, &graph_));
            STFEM_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
          }
        STFEM_HIP(hipGraphLaunch(exec_, ctx_.stream));
      }
    else
      vcycle();
    if constexpr (std::is_same_v<T, double>)
      STFEM_HIP(hipMemcpyAsync(dst, sol_[L].get(), sz * sizeof(double), hipMemcpyDeviceToDevice,
                               ctx_.stream));
    else
      launch_convert<double, T>(sz, sol_[L].get(), dst, ctx_.stream);
  }

  // ================================================================ FGMRES
  SolverFGMRES::SolverFGMRES(Context &ctx, long n, const SolverControlData &c)
    : ctx_(ctx)
    , n_(n)
    , c_(c)
  {
    STFEM_REQUIRE(c.basis >= 2 && c.basis <= 60, "FGMRES: basis size must be in [2, 60]");
    aux_.alloc(n);
  }

  template <typename Op, typename Prec>
  int SolverFGMRES::solve(const Op &A, const Prec &P, double *x, const double *b)
  {
    const int m = c_.basis;
    while ((int)v_.size() < m + 1)
      {
        v_.emplace_back(size_t(n_));
        z_.emplace_back(size_t(n_));
      }
    if (c_.orthogonalization == 1 && gs_G_.size() < size_t(m + 1) * (m + 1))
      {
        gs_part_.alloc(size_t(2 * m + 1) * reduce_slots());
        gs_dots_.alloc(size_t(2 * m + 1));
        gs_h_.alloc(size_t(m + 1));
        gs_G_.alloc(size_t(m + 1) * (m + 1));
      }
    hist_.clear();
    double       *aux = aux_.get();
    double       *dH  = ctx_.scalars.get();
    double       *prt = ctx_.partial.get();
    int           acc = 0;
    double        reduced_tol = 0.0;
    int           state = 0;
    auto          check = [&](int step, double val) {
      hist_.push_back(val);
      if (step == 0)
        reduced_tol = val * c_.reltol;
      last_step_ = step;
      if (!std::isfinite(val))
        return -2;
      if (val <= reduced_tol || val <= c_.abstol)
        return 1;
      if (step >= c_.maxiter)
        return -1;
      return 0;
    };
    std::vector<double> hcol(m + 1);
    // two partial-sum buffers: each Gram-Schmidt step reads the previous
    // step's partials and writes its own
    const int           ns = reduce_slots();
    std::vector<double> hpart(ns);
    double *const       pA = prt, *const pB = prt + ns;
    auto                host_sum = [&](const double *dpart) {
      STFEM_HIP(hipMemcpyAsync(hpart.data(), dpart, ns * sizeof(double), hipMemcpyDeviceToHost,
                               ctx_.stream));
      ctx_.sync();
      return sum_partials(hpart.data());
    };
    do
      {
        A.vmult(aux, x);
        launch_axpby<double>(n_, 1.0, b, -1.0, aux, ctx_.stream);
        launch_dot_partials<double>(n_, aux, aux, pA, ctx_.stream);
        const double beta = std::sqrt(host_sum(pA));
        state = check(acc, beta);
        if (state != 0)
          break;
        std::vector<std::vector<double>> H(m + 1, std::vector<double>(m, 0.0));
        std::vector<double>              y;
        double                           a = beta;
        for (int j = 0; j < m; ++j)
          {
            if (a != 0.0)
              launch_axpby<double>(n_, 1.0 / a, aux, 0.0, v_[j].get(), ctx_.stream);
            else
              v_[j].zero(ctx_.stream);
            P.vmult(z_[j].get(), v_[j].get());
            A.vmult(aux, z_[j].get());
            double *pin = pA;
            if (c_.orthogonalization == 1)
              {
                // modified Gram-Schmidt coefficients from one pass of dots
                // (the Gram matrix correction, all on the device)
                std::vector<const double *> vp(j + 1);
                for (int i = 0; i <= j; ++i)
                  vp[i] = v_[i].get();
                launch_gs_step(n_, j, vp.data(), aux, gs_G_.get(), m + 1, gs_part_.get(),
                               gs_dots_.get(), gs_h_.get(), pA, ctx_.stream);
                STFEM_HIP(hipMemcpyAsync(hcol.data(), gs_h_.get(), (j + 1) * sizeof(double),
                                         hipMemcpyDeviceToHost, ctx_.stream));
              }
            else
              {
                // modified Gram-Schmidt: every step takes its coefficient
                // from the previous step's partial sums (all on the device)
                // and stores it to dH
                launch_dot_partials<double>(n_, aux, v_[0].get(), pA, ctx_.stream);
                double *pout = pB;
                for (int i = 1; i <= j + 1; ++i)
                  {
                    launch_axpy_dot_chain(n_, aux, pin, -1.0, v_[i - 1].get(),
                                          i <= j ? v_[i].get() : nullptr, pout, dH + i - 1,
                                          ctx_.stream);
                    std::swap(pin, pout);
                  }
                STFEM_HIP(hipMemcpyAsync(hcol.data(), dH, (j + 1) * sizeof(double),
                                         hipMemcpyDeviceToHost, ctx_.stream));
              }
            const double hn = std::sqrt(host_sum(pin)); // also waits for hcol
            for (int i = 0; i <= j; ++i)
              H[i][j] = hcol[i];
            H[j + 1][j] = hn;
            a           = hn;
            const int ncol = j; // projected solve once a second vector exists
            if (ncol > 0)
              {
                std::vector<std::vector<double>> R(ncol + 1, std::vector<double>(ncol));
                for (int r0 = 0; r0 <= ncol; ++r0)
                  for (int c0 = 0; c0 < ncol; ++c0)
                    R[r0][c0] = H[r0][c0];
                std::vector<double> g(ncol + 1, 0.0);
                g[0] = beta;
                for (int c0 = 0; c0 < ncol; ++c0)
                  {
                    const double h1 = R[c0][c0], h2 = R[c0 + 1][c0];
                    const double rr = std::hypot(h1, h2);
                    const double cs = rr == 0 ? 1 : h1 / rr, sn = rr == 0 ? 0 : h2 / rr;
                    for (int c1 = c0; c1 < ncol; ++c1)
                      {
                        const double t1 = R[c0][c1], t2 = R[c0 + 1][c1];
                        R[c0][c1]     = cs * t1 + sn * t2;
                        R[c0 + 1][c1] = -sn * t1 + cs * t2;
                      }
                    const double g1 = g[c0], g2 = g[c0 + 1];
                    g[c0]     = cs * g1 + sn * g2;
                    g[c0 + 1] = -sn * g1 + cs * g2;
                  }
                y.assign(ncol, 0.0);
                for (int c0 = ncol - 1; c0 >= 0; --c0)
                  {
                    double s = g[c0];
                    for (int c1 = c0 + 1; c1 < ncol; ++c1)
                      s -= R[c0][c1] * y[c1];
                    y[c0] = s / R[c0][c0];
                  }
                state = check(++acc, std::abs(g[ncol]));
                if (state != 0)
                  break;
              }
          }
        if (!y.empty())
          {
            std::vector<const double *> zp(y.size());
            for (size_t j = 0; j < y.size(); ++j)
              zp[j] = z_[j].get();
            launch_multi_axpy(n_, int(y.size()), zp.data(), y.data(), x, ctx_.stream);
          }
      }
    while (state == 0);
    ctx_.sync();
    if (state == -2)
      return STATUS_ERR_NOT_FINITE;
    return state == 1 ? STATUS_OK : STATUS_ERR_NO_CONVERGE;
  }

  // ================================================================ Problem
  namespace
  {
    template <typename T>
    class Problem final : public ProblemBase
    {
    public:
      Problem(const ProblemSpec &s, const double *vertices, const double *kappa, int device)
        : ctx_(device)
        , sh_(s.p)
      {
        spec = s;
        if (spec.print_timing)
          ctx_.timers = std::make_unique<PhaseTimers>();
        PhaseLog plog;
        mesh_ = build_mesh_levels(spec.dim, spec.subdiv, spec.refine, spec.p, vertices, spec.lo,
                                  spec.hi, kappa, spec.dirichlet, spec.coarse_coef);
        plog("mesh levels");
        // time step (tp_01.cc:87, 124-127) unless given
        if (spec.tau > 0)
          tau = spec.tau;
        else
          {
            double diag = 0;
            for (int d = 0; d < spec.dim; ++d)
              {
                const double h = (spec.hi[d] - spec.lo[d]) / spec.subdiv[d];
                diag += h * h;
              }
            const double   spc    = std::sqrt(diag) / std::sqrt(double(spec.dim));
            const unsigned nsteps = unsigned(spec.end_time / spc);
            STFEM_REQUIRE(nsteps >= 1, "spec: end time shorter than the coarse step");
            tau = spec.end_time * std::pow(2.0, -(spec.refine + 1)) / nsteps;
          }
        gmg_ = build_gmg(false, plog);
        plog("relaxation estimates and dense coarse operator");
        // FP64 fine system and rhs matrices (tp_01.cc:152-198, heat)
        const MeshLevel  &F  = mesh_.back();
        const TimeWeights w  = get_fe_time_weights(spec.type, spec.r, tau, spec.nts);
        const TimeWeights w1 = get_fe_time_weights(spec.type, spec.r, tau, 1);
        const bool        tp = spec.atomic_scatter == 0;
        nt                   = w.Alpha.m;
        DMat zero(w.Gamma.m, 1);
        if (spec.problem == 1)
          {
            // tp_01.cc:167-197: system (Alpha_lhs, Beta_lhs), rhs_matrix
            // (rhs_uK, rhs_uM) on prev_u, rhs_matrix_v (0, rhs_vM) on prev_v
            const WaveWeights ww =
              get_fe_time_weights_wave(spec.type, w1.Alpha, w1.Beta, w1.Gamma, w1.Zeta, spec.nts);
            fine_ = std::make_unique<SystemMatrix<double>>(ctx_, F, ww.lhs_uK, ww.lhs_uM, tp);
            rhs_  = std::make_unique<SystemMatrix<double>>(ctx_, F, ww.rhs_uK, ww.rhs_uM, tp);
            rhsv_ = std::make_unique<SystemMatrix<double>>(ctx_, F, DMat(ww.rhs_vM.m, 1),
                                                           ww.rhs_vM, tp);
            rhsv_->set_max_jb(spec.op_jb64);
            rhsv_->set_split_below(spec.op_split);
            rhsv_->set_op_kernel(spec.op_kernel);
            // TimeIntegratorWave velocity matrices (time_integrators.h:219-233)
            const DMat Ai = w1.Alpha.inverse();
            DMat       AixB = Ai * w1.Beta, AixG = Ai * w1.Gamma, AixZ = Ai * w1.Zeta;
            if (spec.type == TimeStepType::DG)
              AixG.scale(-1.0);
            else
              AixZ.scale(-1.0);
            aixb_.upload(AixB.a.data(), AixB.a.size(), ctx_.stream);
            aixg_.upload(AixG.a.data(), AixG.a.size(), ctx_.stream);
            aixz_.upload(AixZ.a.data(), AixZ.a.size(), ctx_.stream);
          }
        else
          {
            fine_ = std::make_unique<SystemMatrix<double>>(ctx_, F, w.Alpha, w.Beta, tp);
            const DMat rK = (spec.type == TimeStepType::CGP) ? w.Gamma : zero;
            const DMat rM = (spec.type == TimeStepType::CGP) ? w.Zeta : w.Gamma;
            rhs_ = std::make_unique<SystemMatrix<double>>(ctx_, F, rK, rM, tp);
          }
        // every FP64 operator with the same output chunking
        for (auto *op : {fine_.get(), rhs_.get()})
          {
            op->set_max_jb(spec.op_jb64);
            op->set_split_below(spec.op_split);
            op->set_op_kernel(spec.op_kernel);
          }
        alpha1_ = w1.Alpha;
        gamma1_ = w1.Gamma;
        plog("FP64 system and rhs operators");
        solver_ = std::make_unique<SolverFGMRES>(ctx_, fine_->m(), spec.solver);
        plog("FGMRES basis");
        rhs_vec_.alloc(size_t(fine_->m()));
        coefs_.alloc(64);
        ones_.alloc(64);
        {
          std::vector<double> o(64, 1.0);
          ones_.upload(o.data(), 64, ctx_.stream);
        }
        if (spec.source == 0)
          build_load([this](const double *x) {
            const double PI = 3.14159265358979323846, f = spec.frequency;
            double       hv = 1.0;
            for (int d = 0; d < spec.dim; ++d)
              hv *= std::sin(2 * PI * f * x[d]);
            return hv;
          });
        else if (spec.source == 3)
          {
            // seeded Gaussian bumps (spec.h: gaussian_bumps)
            const std::vector<double> bumps = gaussian_bumps(spec.n_bumps, spec.seed);
            build_load(
              [&](const double *x) { return gaussian_bumps_value(bumps, spec.dim, x); });
          }
        ctx_.sync();
        plog("load vector");
      }

      // The space-time multigrid of tp_01.cc:200-359 / GMG::reinit
      // (gmg.h:708-790): level sequence, level operators, Vanka smoothers,
      // transfers.  transpose: the same hierarchy for the transposed system
      // (time matrices transposed on every level), the preconditioner of the
      // adjoint solves.
      std::unique_ptr<GMG<T>> build_gmg(bool transpose, PhaseLog &plog)
      {
        // level sequence (tp_01.cc:205-219)
        const int n_sp = spec.mg_space ? spec.refine + 1 : 1;
        const int rmin = spec.space_time_mg ? spec.r_min : spec.r;
        const int tmin = spec.space_time_mg ? std::max(spec.nts_min, 1) : spec.nts;
        const auto seq = get_time_mg_sequence(n_sp, spec.r, rmin, spec.nts, tmin,
                                              spec.time_before_space != 0);
        const int nlev = int(seq.size()) + 1;
        std::vector<int>    sref(nlev), rl(nlev), nl(nlev);
        std::vector<double> tl(nlev);
        sref[nlev - 1] = spec.refine;
        rl[nlev - 1]   = spec.r;
        nl[nlev - 1]   = spec.nts;
        tl[nlev - 1]   = tau;
        for (int l = nlev - 2; l >= 0; --l)
          {
            const TimeMGType t = seq[l];
            sref[l] = sref[l + 1] - ((t == TimeMGType::none && spec.mg_space) ? 1 : 0);
            rl[l]   = rl[l + 1] - (t == TimeMGType::k ? 1 : 0);
            nl[l]   = nl[l + 1] / (t == TimeMGType::tau ? 2 : 1);
            tl[l]   = tl[l + 1] * (t == TimeMGType::tau ? 2.0 : 1.0);
          }
        std::vector<GMGLevel<T>> levels(nlev);
        const auto form = spec.mg.vanka_factored ? PreconditionVanka<T>::Form::factored :
                                                   PreconditionVanka<T>::Form::explicit_inverse;
        for (int l = 0; l < nlev; ++l)
          {
            TimeWeights w = get_fe_time_weights(spec.type, rl[l], tl[l], nl[l]);
            if (transpose)
              {
                // adjoint levels: (Alpha^T x K + Beta^T x M), K and M symmetric
                w.Alpha = w.Alpha.transposed();
                w.Beta  = w.Beta.transposed();
              }
            if (spec.problem == 1)
              {
                // wave levels: the Schur-in-time transform of the level's
                // multi-step weights taken as one interval (fe_time.h:383-412)
                const WaveWeights ww =
                  get_fe_time_weights_wave(spec.type, w.Alpha, w.Beta, w.Gamma, w.Zeta, 1);
                w.Alpha = ww.lhs_uK;
                w.Beta  = ww.lhs_uM;
              }
            const MeshLevel &m = mesh_[sref[l]];
            levels[l].A  = std::make_unique<SystemMatrix<T>>(ctx_, m, w.Alpha, w.Beta,
                                                             spec.atomic_scatter == 0);
            levels[l].A->set_max_jb(sizeof(T) == 4 ? spec.op_jb : spec.op_jb64);
            levels[l].A->set_split_below(spec.op_split);
            levels[l].A->set_op_kernel(spec.op_kernel);
            levels[l].V  = std::make_unique<PreconditionVanka<T>>(ctx_, m, sh_, w.Alpha, w.Beta, form,
                                                                  spec.vanka_device_setup);
            levels[l].V->set_variant(spec.vanka_kernel);
            levels[l].r  = rl[l];
            levels[l].nts = nl[l];
            levels[l].type_to_here = l > 0 ? seq[l - 1] : TimeMGType::none;
            ctx_.sync();
            plog("level operator and Vanka", l);
          }
        std::vector<TwoLevelTransfer<T>> tr(nlev);
        for (int l = 1; l < nlev; ++l)
          {
            const TimeMGType t = seq[l - 1];
            if (t == TimeMGType::none)
              tr[l].space = std::make_unique<MGTwoLevelBlockTransfer<T>>(
                ctx_, mesh_[sref[l]].g, mesh_[sref[l - 1]].g, sh_, levels[l].A->n_blocks());
            else
              tr[l].time = std::make_unique<MGTwoLevelTransferST<T>>(
                ctx_, spec.type, rl[l], nl[l], t, mesh_[sref[l]].g.n_nodes,
                spec.mg.overwrite_time_transfer, spec.restrict_is_transpose != 0);
          }
        plog("transfers");
        return std::make_unique<GMG<T>>(ctx_, std::move(levels), std::move(tr), spec.mg);
      }

      int n_levels() const override { return gmg_->n_levels(); }
      void level_info(int l, long *info) const override
      {
        const auto &L = gmg_->level(l);
        const auto &g = L.A->grid();
        info[0] = g.n_nodes;
        info[1] = L.A->n_blocks();
        info[2] = g.n_cells;
        info[3] = L.r;
        info[4] = L.nts;
        info[5] = l > 0 ? int(L.type_to_here) : 0;
        info[6] = g.nc[0];
        info[7] = g.nc[1];
        info[8] = g.nc[2];
        info[9] = sizeof(T);
        info[10] = l <= gmg_->dense_level() ? 1 : 0;
        info[11] = l == gmg_->dense_level() ? g.n_nodes * L.A->n_blocks() : 0;
      }
      double level_omega(int l) const override { return gmg_->level(l).omega; }
      void level_vmult(int l, void *dst, const void *src, bool transpose) override
      {
        const auto &L = gmg_->level(l);
        if (transpose)
          L.A->Tvmult((T *)dst, (const T *)src);
        else
          L.A->vmult((T *)dst, (const T *)src);
      }
      void level_vanka(int l, void *dst, const void *src) override
      {
        gmg_->level(l).V->vmult((T *)dst, (const T *)src);
      }
      void level_vanka_cells(int l, const void *src) override
      {
        gmg_->level(l).V->cell_pass((const T *)src);
      }
      void level_prolongate(int l, void *dst, const void *src) override
      {
        STFEM_REQUIRE(l >= 1 && l < n_levels(), "prolongate: bad level");
        gmg_->transfer(l).prolongate_and_add((T *)dst, (const T *)src);
      }
      void level_restrict(int l, void *dst, const void *src) override
      {
        STFEM_REQUIRE(l >= 1 && l < n_levels(), "restrict: bad level");
        gmg_->transfer(l).restrict_and_add((T *)dst, (const T *)src);
      }
      void fine_vmult(double *dst, const double *src, bool transpose) override
      {
        transpose ? fine_->Tvmult(dst, src) : fine_->vmult(dst, src);
      }
      void gmg_vmult(double *dst, const double *src) override { gmg_->vmult(dst, src); }
      int  fgmres(double *x, const double *b, int *its, std::vector<double> *hist) override
      {
        const int rc = solver_->solve(*fine_, *gmg_, x, b);
        if (its)
          *its = solver_->last_step();
        if (hist)
          *hist = solver_->history();
        return rc;
      }
      // TimeIntegratorHeat::solve (time_integrators.h:157-180)
      int solve_slab(double *x, const double *prev, double time, int *its) override
      {
        const PhaseScope scope(ctx_, PhaseTimers::step);
        STFEM_REQUIRE(spec.problem == 0, "solve_slab: the wave problem needs solve_slab_wave");
        return solve_displacement(x, prev, nullptr, time, its);
      }
      // TimeIntegratorWave::solve (time_integrators.h:237-280): the
      // displacement slab with rhs_matrix_v * prev_v added to the right-hand
      // side, then the velocities of every step from the slab's
      // displacements and the previous step's last displacement/velocity
      int solve_slab_wave(double *x, double *v, const double *prev_u, const double *prev_v,
                          double time, int *its) override
      {
        const PhaseScope scope(ctx_, PhaseTimers::step);
        STFEM_REQUIRE(spec.problem == 1, "solve_slab_wave: the problem is not the wave problem");
        const int rc = solve_displacement(x, prev_u, prev_v, time, its);
        launch_wave_velocity(fine_->n_nodes(), spec.nts, nt / spec.nts,
                             spec.type == TimeStepType::DG, aixb_.get(), aixg_.get(),
                             aixz_.get(), x, prev_u, prev_v, v, ctx_.stream);
        return rc;
      }
      int solve_displacement(double *x, const double *prev, const double *prev_v, double time,
                             int *its)
      {
        const long N  = fine_->n_nodes();
        double    *rh = rhs_vec_.get();
        STFEM_HIP(hipMemsetAsync(rh, 0, size_t(N) * nt * sizeof(double), ctx_.stream));
        if (prev_v)
          rhsv_->vmult_add_nx1(rh, prev_v);
        rhs_->vmult_add_nx1(rh, prev);
        if (spec.source != 1 && load_.size())
          {
            // assemble_force (time_integrators.h:60-87): source = g(t) * h(x)
            const LagrangeBasis tb = get_time_basis(spec.type, spec.r);
            const int           ni = alpha1_.m;
            std::vector<double> c(nt, 0.0);
            for (int it = 0; it < spec.nts; ++it)
              for (int j = 0; j < tb.size(); ++j)
                {
                  const double t = time + tau * it + tau * tb.nodes[j];
                  const double g = source_time(t);
                  if (spec.type == TimeStepType::DG)
                    c[j + it * ni] += alpha1_(j, j) * g;
                  else if (j == 0)
                    for (int i = 0; i < gamma1_.m; ++i)
                      c[i + it * ni] += -gamma1_(i, 0) * g;
                  else
                    c[j - 1 + it * ni] += alpha1_(j - 1, j - 1) * g;
                }
            coefs_.upload(c.data(), c.size(), ctx_.stream);
            launch_time_transfer<double>(N, coefs_.get(), nt, 1, load_.get(), rh, true,
                                         ctx_.stream);
          }
        // start vector (time_integrators.h:96-104): with extrapolation every
        // block starts from prev, otherwise from zero
        if (spec.extrapolate)
          launch_time_transfer<double>(N, ones_.get(), nt, 1, prev, x, false, ctx_.stream);
        else
          STFEM_HIP(hipMemsetAsync(x, 0, size_t(N) * nt * sizeof(double), ctx_.stream));
        const int rc = solver_->solve(*fine_, *gmg_, x, rh);
        launch_zero_dirichlet<double>(fine_->grid(), nt, x, ctx_.stream);
        if (its)
          *its = solver_->last_step();
        return rc;
      }
      void slab_error(const double *x, const double *prev, double time, double *out) override
      {
        STFEM_REQUIRE(spec.source == 0,
                      "slab error: needs the manufactured solution (source=0)");
        const GridDesc &g  = fine_->grid();
        const int       nq = spec.r + 1, n = sh_.n;
        const Quad1D    sq = quad_gauss(nq);
        std::vector<double> S(size_t(nq) * n), D(size_t(nq) * n);
        for (int q = 0; q < nq; ++q)
          for (int a = 0; a < n; ++a)
            {
              S[q * n + a] = sh_.L.value(a, sq.x[q]);
              D[q * n + a] = sh_.L.deriv(a, sq.x[q]);
            }
        // time points: the numerical solution is a combination of the
        // slab's blocks and, for cG, of the previous slab's last block
        const LagrangeBasis tb = get_time_basis(spec.type, spec.r);
        const Quad1D        tq = quad_gauss(spec.r + 1);
        const int           ni = n_dofs_per_interval(spec.type, spec.r);
        std::vector<double> rows;
        for (int it = 0; it < spec.nts; ++it)
          for (size_t q = 0; q < tq.x.size(); ++q)
            {
              std::vector<double> row(size_t(nt) + 3, 0.0);
              row[0] = time + tau * it + tq.x[q] * tau;
              row[1] = tau * tq.w[q];
              for (int i = 0; i < tb.size(); ++i)
                {
                  const double v = tb.value(i, tq.x[q]);
                  int          k;
                  if (spec.type == TimeStepType::DG)
                    k = it * ni + i;
                  else
                    k = it * ni + i - 1; // -1: the previous slab's last block
                  row[2 + (k < 0 ? nt : k)] += v;
                }
              rows.insert(rows.end(), row.begin(), row.end());
            }
        const int npts = int(rows.size() / (size_t(nt) + 3));
        err_tp_.upload(rows.data(), rows.size(), ctx_.stream);
        if (err_cells_.size() < size_t(g.n_cells) * 3)
          err_cells_.alloc(size_t(g.n_cells) * 3);
        launch_slab_error(g, fine_->vertices(), nt, x, prev, nq, sq.x.data(), sq.w.data(),
                          S.data(), D.data(), npts, err_tp_.get(), spec.frequency,
                          err_cells_.get(), ctx_.stream);
        std::vector<double> e(size_t(g.n_cells) * 3);
        STFEM_HIP(hipMemcpyAsync(e.data(), err_cells_.get(), e.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx_.stream));
        ctx_.sync();
        out[0] = 0;
        out[1] = -1;
        out[2] = 0;
        for (long k = 0; k < g.n_cells; ++k)
          {
            out[0] += e[k * 3];
            out[1] = std::max(out[1], e[k * 3 + 1]);
            out[2] += e[k * 3 + 2];
          }
      }
      int run_heat(double *stats, int *its_per_slab, int cap) override
      {
        const long   N    = fine_->n_nodes();
        const bool   wave = spec.problem == 1;
        DBuf<double> x{size_t(N) * nt}, prev{size_t(N)}, last{size_t(nt)};
        DBuf<double> v{wave ? size_t(N) * nt : 0}, prev_v{wave ? size_t(N) : 0};
        launch_fill_zero<double>(x.get(), x.size(), ctx_.stream);
        std::vector<double> sel(nt, 0.0);
        sel[nt - 1] = 1.0; // picks the last block of a node-major vector
        last.upload(sel.data(), sel.size(), ctx_.stream);
        // initial values (tp_01.cc:566-568): u(0) = 0 for the manufactured
        // solution, the wave velocity interpolates v(0) = 2 pi f prod sin
        if (wave)
          {
            const std::vector<double> v0 = interpolate_exact_v(0.0);
            prev_v.upload(v0.data(), size_t(N), ctx_.stream);
            ctx_.sync();
          }
        double time = 0.0, l2 = 0.0, l8 = -1.0, h1 = 0.0;
        int    slab = 0, total = 0, rc = 0;
        while (time < spec.end_time)
          {
            launch_time_transfer<double>(N, last.get(), 1, nt, x.get(), prev.get(), false,
                                         ctx_.stream);
            if (wave && slab > 0)
              launch_time_transfer<double>(N, last.get(), 1, nt, v.get(), prev_v.get(), false,
                                           ctx_.stream);
            int       its = 0;
            const int r   = wave ? solve_slab_wave(x.get(), v.get(), prev.get(), prev_v.get(),
                                                   time, &its)
                                 : solve_slab(x.get(), prev.get(), time, &its);
            if (r != 0)
              rc = r;
            if (its_per_slab && slab < cap)
              its_per_slab[slab] = its;
            total += its;
            double e[3];
            slab_error(x.get(), prev.get(), time, e);
            l2 += e[0];
            l8 = std::max(l8, e[1]);
            h1 += e[2];
            time += spec.nts * tau;
            ++slab;
          }
        stats[0] = slab;
        stats[1] = total;
        stats[2] = std::sqrt(l2);
        stats[3] = l8;
        stats[4] = std::sqrt(h1);
        stats[5] = double(N);
        stats[6] = nt;
        return rc;
      }
      // ------------------------------------------------------------ adjoint
      // Discrete adjoint of the heat slab sequence A x_n = R E x_{n-1} + f_n
      // (TimeIntegratorHeat, time_integrators.h:157-180), x_{-1} = 0, with
      // the misfit J = 1/2 sum_n e_n^T (Alpha (x) M) e_n, e_n = x_n - d_n
      // (for dG, Alpha is the time mass matrix, so J is half the squared
      // L2(space-time) norm of the mismatch).  Backward in n:
      //   A^T lambda_n = (Alpha (x) M) e_n + E^T R^T lambda_{n+1}
      //   dJ/dtheta_c = -sum_n lambda_n^T (dA/dtheta_c) x_n
      // (R = Gamma (x) M carries no coefficient for dG).  Only the slab start
      // states are kept from the forward sweep; each slab is solved again
      // before its adjoint solve.  The adjoint solves use FGMRES with the
      // multigrid of the transposed system.
      void ensure_adjoint()
      {
        STFEM_REQUIRE(spec.problem == 0, "adjoint: heat problem only");
        STFEM_REQUIRE(spec.type == TimeStepType::DG, "adjoint: dG time discretisation only");
        if (gmg_adj_)
          return;
        const MeshLevel &F = mesh_.back();
        const bool       tp = spec.atomic_scatter == 0;
        const DMat      &A = fine_->alpha();
        fineT_  = std::make_unique<SystemMatrix<double>>(ctx_, F, A.transposed(),
                                                         fine_->beta().transposed(), tp);
        misfit_ = std::make_unique<SystemMatrix<double>>(ctx_, F, DMat(A.m, A.n), A, tp);
        rhsT_   = std::make_unique<SystemMatrix<double>>(ctx_, F, rhs_->alpha().transposed(),
                                                         rhs_->beta().transposed(), tp);
        dalpha_.upload(A.a.data(), A.a.size(), ctx_.stream);
        std::vector<double> k(size_t(F.g.n_cells), F.coef_mode == 0 ? F.coef_value : 0.0);
        if (F.coef_mode == 1)
          k.assign(F.coef.begin(), F.coef.end());
        kappa_.upload(k.data(), k.size(), ctx_.stream);
        std::vector<double> sel(nt, 0.0);
        sel[nt - 1] = 1.0;
        last_.upload(sel.data(), sel.size(), ctx_.stream); // picks / places the last block
        ctx_.sync();
        PhaseLog plog;
        gmg_adj_ = build_gmg(true, plog);
        plog("adjoint multigrid");
      }
      // 1/2 e^T (Alpha x M) e with e = x - d (e, qe: scratch)
      double slab_misfit(const double *x, const double *d, double *e, double *qe)
      {
        const long m = fine_->m();
        STFEM_HIP(hipMemcpyAsync(e, x, size_t(m) * sizeof(double), hipMemcpyDeviceToDevice,
                                 ctx_.stream));
        launch_axpby<double>(m, -1.0, d, 1.0, e, ctx_.stream);
        misfit_->vmult(qe, e);
        launch_dot_partials<double>(m, e, qe, ctx_.partial.get(), ctx_.stream);
        std::vector<double> h(static_cast<size_t>(reduce_slots()));
        STFEM_HIP(hipMemcpyAsync(h.data(), ctx_.partial.get(), h.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx_.stream));
        ctx_.sync();
        return 0.5 * sum_partials(h.data());
      }
      void forward_trajectory(int n_slabs, double *traj, int *its) override
      {
        STFEM_REQUIRE(n_slabs >= 1, "trajectory: at least one slab");
        const long   N = fine_->n_nodes(), m = fine_->m();
        DBuf<double> x{size_t(m)}, prev{size_t(N)};
        std::vector<double> sel(nt, 0.0);
        sel[nt - 1] = 1.0;
        DBuf<double> last;
        last.upload(sel.data(), sel.size(), ctx_.stream);
        launch_fill_zero<double>(x.get(), x.size(), ctx_.stream);
        for (int n = 0; n < n_slabs; ++n)
          {
            launch_time_transfer<double>(N, last.get(), 1, nt, x.get(), prev.get(), false,
                                         ctx_.stream);
            int       k = 0;
            const int r = solve_slab(x.get(), prev.get(), n * spec.nts * tau, &k);
            if (its)
              its[n] = k;
            if (r != 0)
              throw Error(r, "forward trajectory: slab solve failed");
            STFEM_HIP(hipMemcpyAsync(traj + size_t(n) * m, x.get(), size_t(m) * sizeof(double),
                                     hipMemcpyDeviceToDevice, ctx_.stream));
          }
        ctx_.sync();
      }
      double misfit(const double *obs, int n_slabs, int *its) override
      {
        ensure_adjoint();
        const long   N = fine_->n_nodes(), m = fine_->m();
        DBuf<double> x{size_t(m)}, prev{size_t(N)}, e{size_t(m)}, qe{size_t(m)};
        launch_fill_zero<double>(x.get(), x.size(), ctx_.stream);
        double J = 0.0;
        for (int n = 0; n < n_slabs; ++n)
          {
            launch_time_transfer<double>(N, last_.get(), 1, nt, x.get(), prev.get(), false,
                                         ctx_.stream);
            int       k = 0;
            const int r = solve_slab(x.get(), prev.get(), n * spec.nts * tau, &k);
            if (its)
              its[n] = k;
            if (r != 0)
              throw Error(r, "misfit: slab solve failed");
            J += slab_misfit(x.get(), obs + size_t(n) * m, e.get(), qe.get());
          }
        return J;
      }
      void adjoint_gradient(const double *obs, int n_slabs, double *grad, double *J,
                            int *its) override
      {
        STFEM_REQUIRE(n_slabs >= 1, "adjoint: at least one slab");
        ensure_adjoint();
        const GridDesc &g = fine_->grid();
        const long      N = g.n_nodes, m = fine_->m();
        DBuf<double> ck{size_t(n_slabs) * N}, x{size_t(m)}, e{size_t(m)}, qe{size_t(m)},
          lam{size_t(m)}, w{size_t(m)}, lr{size_t(N)};
        // forward sweep: slab start states only
        launch_fill_zero<double>(x.get(), x.size(), ctx_.stream);
        double Jv = 0.0;
        for (int n = 0; n < n_slabs; ++n)
          {
            double *prev = ck.get() + size_t(n) * N;
            launch_time_transfer<double>(N, last_.get(), 1, nt, x.get(), prev, false,
                                         ctx_.stream);
            int       k = 0;
            const int r = solve_slab(x.get(), prev, n * spec.nts * tau, &k);
            if (its)
              its[2 * n] = k;
            if (r != 0)
              throw Error(r, "adjoint: forward slab solve failed");
            Jv += slab_misfit(x.get(), obs + size_t(n) * m, e.get(), qe.get());
          }
        // backward sweep
        launch_fill_zero<double>(grad, size_t(g.n_cells), ctx_.stream);
        for (int n = n_slabs - 1; n >= 0; --n)
          {
            const double *prev = ck.get() + size_t(n) * N;
            int           k    = 0;
            int           r    = solve_slab(x.get(), prev, n * spec.nts * tau, &k);
            if (r != 0)
              throw Error(r, "adjoint: slab recomputation failed");
            // adjoint right-hand side: (Alpha x M) e_n + E^T R^T lambda_{n+1}
            (void)slab_misfit(x.get(), obs + size_t(n) * m, e.get(), qe.get());
            if (n + 1 < n_slabs)
              launch_time_transfer<double>(N, last_.get(), nt, 1, lr.get(), qe.get(), true,
                                           ctx_.stream);
            launch_fill_zero<double>(lam.get(), lam.size(), ctx_.stream);
            r = solver_->solve(*fineT_, *gmg_adj_, lam.get(), qe.get());
            if (its)
              its[2 * n + 1] = solver_->last_step();
            if (r != 0)
              throw Error(r, "adjoint: adjoint slab solve failed");
            launch_zero_dirichlet<double>(g, nt, lam.get(), ctx_.stream);
            // g_c -= kappa_c sum_j lambda_j^T K_c (Alpha x)_j
            launch_time_transfer<double>(N, dalpha_.get(), nt, nt, x.get(), w.get(), false,
                                         ctx_.stream);
            launch_cell_grad(g, fine_->vertices(), nt, lam.get(), w.get(), kappa_.get(),
                             sh_.qx.data(), sh_.qw.data(), sh_.S.data(), sh_.Sd.data(), grad,
                             ctx_.stream);
            // R^T lambda_n for the previous slab's last block
            rhsT_->vmult(lr.get(), lam.get());
          }
        ctx_.sync();
        *J = Jv;
      }
      void set_cell_coefficient(const double *kappa) override
      {
        MeshLevel &F = mesh_.back();
        F.coef_mode  = 1;
        F.coef.assign(kappa, kappa + F.g.n_cells);
        fine_->set_cell_coefficient(kappa);
        rhs_->set_cell_coefficient(kappa);
        if (rhsv_)
          rhsv_->set_cell_coefficient(kappa);
        if (gmg_adj_)
          {
            fineT_->set_cell_coefficient(kappa);
            rhsT_->set_cell_coefficient(kappa);
            kappa_.upload(kappa, size_t(F.g.n_cells), ctx_.stream);
            ctx_.sync();
          }
      }

      void set_load_vector(const double *h) override
      {
        load_.upload(h, size_t(fine_->n_nodes()), ctx_.stream);
        ctx_.sync();
      }
      long     n_nodes() const override { return fine_->n_nodes(); }
      Context &context() override { return ctx_; }
      double   fine_vmult_bytes() const override { return fine_->vmult_bytes(); }
      double   vanka_sweep_bytes(int l) const override
      {
        return gmg_->level(l).V->sweep_bytes();
      }
      const std::vector<double> &history() const { return solver_->history(); }

    private:
      DBuf<double> err_tp_, err_cells_; // slab error: time table, per-cell errors
      // VectorTools::interpolate of wave::ExactSolutionV (exact_solution.h:
      // 143-163) at the fine nodes: support points are the GLL points mapped
      // by each cell's Q1 map; constrained nodes stay 0
      std::vector<double> interpolate_exact_v(double t) const
      {
        const GridDesc     &g = mesh_.back().g;
        const int           dim = g.dim;
        const double        PI = 3.14159265358979323846, f = spec.frequency;
        std::vector<double> xyz;
        node_coordinates_host(xyz);
        std::vector<double> out(g.n_nodes, 0.0);
        for (long k = 0; k < g.n_nodes; ++k)
          {
            const long gx = k % g.nn[0], gy = (k / g.nn[0]) % g.nn[1],
                       gz = dim == 3 ? k / (long(g.nn[0]) * g.nn[1]) : 0;
            if (node_constrained(g, gx, gy, gz))
              continue;
            double val = 2 * PI * f * std::cos(2 * PI * f * t);
            for (int d = 0; d < dim; ++d)
              val *= std::sin(2 * PI * f * xyz[k * dim + d]);
            out[k] = val;
          }
        return out;
      }

    public:
      // support points of the fine FE_Q(p) nodes: Gauss-Lobatto points mapped
      // by each cell's Q1 map (DoFTools::map_dofs_to_support_points); a node
      // shared by several cells takes the value of the last cell visited
      void node_coordinates_host(std::vector<double> &xyz) const override
      {
        const MeshLevel &m = mesh_.back();
        const GridDesc  &g = m.g;
        const int        dim = g.dim, n = sh_.n;
        const long       nv0 = g.nc[0] + 1, nv1 = g.nc[1] + 1;
        xyz.assign(size_t(g.n_nodes) * dim, 0.0);
        for (long c = 0; c < g.n_cells; ++c)
          {
            int cx, cy, cz;
            cell_xyz(g, c, cx, cy, cz);
            for (int c3 = 0; c3 < ((dim == 3) ? n : 1); ++c3)
              for (int b = 0; b < n; ++b)
                for (int a = 0; a < n; ++a)
                  {
                    const long gx = (long)cx * sh_.p + a, gy = (long)cy * sh_.p + b,
                               gz = (long)cz * sh_.p + c3;
                    const double xi[3] = {sh_.gll[a], sh_.gll[b], dim == 3 ? sh_.gll[c3] : 0.0};
                    double       x[3] = {0, 0, 0};
                    for (int vtx = 0; vtx < (1 << dim); ++vtx)
                      {
                        const long id = (cx + (vtx & 1)) +
                                        nv0 * ((cy + ((vtx >> 1) & 1)) + nv1 * (cz + ((vtx >> 2) & 1)));
                        double N0 = 1;
                        for (int e = 0; e < dim; ++e)
                          N0 *= ((vtx >> e) & 1) ? xi[e] : 1 - xi[e];
                        for (int d = 0; d < dim; ++d)
                          x[d] += N0 * m.X[id * dim + d];
                      }
                    const long k = gx + g.nn[0] * (gy + g.nn[1] * gz);
                    for (int d = 0; d < dim; ++d)
                      xyz[k * dim + d] = x[d];
                  }
          }
      }

    private:
      // time profile g(t) of the source f(t,x) = g(t) h(x): the manufactured
      // heat source of exact_solution.h:53-75 splits this way; the other
      // sources are constant in time
      double source_time(double t) const
      {
        if (spec.source != 0)
          return 1.0;
        const double PI = 3.14159265358979323846, f = spec.frequency;
        if (spec.problem == 1) // wave::RHSFunction (exact_solution.h:166-180)
          return std::pow(2.0, spec.dim) * (PI * f) * (PI * f) * std::sin(2 * PI * f * t);
        return spec.dim * 4 * PI * PI * f * f * std::sin(2 * PI * f * t) +
               2 * PI * f * std::cos(2 * PI * f * t);
      }
      // load vector h_i = int h(x) phi_i with Gauss(p+1) quadrature on the
      // mapped cells (VectorTools::create_right_hand_side), Dirichlet rows 0.
      // Per cell the quadrature values w_q h(x_q) are contracted with the 1D
      // basis one direction at a time (sum factorization); cells run in
      // parallel within each of the 2^dim parity classes of (cx, cy, cz),
      // whose cells share no node, so every node sums its cells in the
      // same order for any thread count.
      template <typename F>
      void build_load(F hfun)
      {
        const MeshLevel &m   = mesh_.back();
        const GridDesc  &g   = m.g;
        const int        dim = g.dim, n = sh_.n;
        const int        nz  = dim == 3 ? n : 1;
        std::vector<double> h(g.n_nodes, 0.0);
        const long          nv0 = g.nc[0] + 1, nv1 = g.nc[1] + 1;
        const long          hc[3] = {(g.nc[0] + 1) / 2, (g.nc[1] + 1) / 2,
                                     dim == 3 ? (g.nc[2] + 1) / 2 : 1};
        for (int color = 0; color < (1 << dim); ++color)
          {
            const int  ox = color & 1, oy = (color >> 1) & 1, oz = (color >> 2) & 1;
            const long ncol = hc[0] * hc[1] * hc[2];
#pragma omp parallel
            {
              std::vector<double> fq(size_t(n) * n * nz), t1(fq.size()), t2(fq.size());
#pragma omp for schedule(static)
              for (long k = 0; k < ncol; ++k)
                {
                  const int cx = int(2 * (k % hc[0]) + ox), cy = int(2 * ((k / hc[0]) % hc[1]) + oy);
                  const int cz = dim == 3 ? int(2 * (k / (hc[0] * hc[1])) + oz) : 0;
                  if (cx >= g.nc[0] || cy >= g.nc[1] || (dim == 3 && cz >= g.nc[2]))
                    continue;
                  double V[8][3] = {};
                  for (int v = 0; v < (1 << dim); ++v)
                    {
                      const long id = (cx + (v & 1)) +
                                      nv0 * ((cy + ((v >> 1) & 1)) + nv1 * (cz + ((v >> 2) & 1)));
                      for (int d = 0; d < dim; ++d)
                        V[v][d] = m.X[id * dim + d];
                    }
                  // fq[qz][qy][qx] = w_q |det J| h(x_q)
                  for (int qz = 0; qz < nz; ++qz)
                    for (int qy = 0; qy < n; ++qy)
                      for (int qx = 0; qx < n; ++qx)
                        {
                          const double xi[3] = {sh_.qx[qx], sh_.qx[qy], dim == 3 ? sh_.qx[qz] : 0.0};
                          double       x[3] = {0, 0, 0}, J[3][3] = {};
                          for (int v = 0; v < (1 << dim); ++v)
                            {
                              double N0 = 1;
                              for (int e = 0; e < dim; ++e)
                                N0 *= ((v >> e) & 1) ? xi[e] : 1 - xi[e];
                              for (int e = 0; e < dim; ++e)
                                {
                                  double gg = 1;
                                  for (int ff = 0; ff < dim; ++ff)
                                    {
                                      const int bit = (v >> ff) & 1;
                                      gg *= (ff == e) ? (bit ? 1.0 : -1.0) : (bit ? xi[ff] : 1 - xi[ff]);
                                    }
                                  for (int d = 0; d < dim; ++d)
                                    J[d][e] += gg * V[v][d];
                                }
                              for (int d = 0; d < dim; ++d)
                                x[d] += N0 * V[v][d];
                            }
                          double det;
                          if (dim == 2)
                            det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
                          else
                            det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                  J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                  J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                          const double w =
                            sh_.qw[qx] * sh_.qw[qy] * (dim == 3 ? sh_.qw[qz] : 1.0) * std::abs(det);
                          fq[(size_t(qz) * n + qy) * n + qx] = hfun(x) * w;
                        }
                  // contract x, then y, then z: t[.][.][a] = sum_q S[q][a] f[.][.][q]
                  for (int qz = 0; qz < nz; ++qz)
                    for (int qy = 0; qy < n; ++qy)
                      for (int a = 0; a < n; ++a)
                        {
                          double acc = 0;
                          for (int qx = 0; qx < n; ++qx)
                            acc += sh_.S[qx * n + a] * fq[(size_t(qz) * n + qy) * n + qx];
                          t1[(size_t(qz) * n + qy) * n + a] = acc;
                        }
                  for (int qz = 0; qz < nz; ++qz)
                    for (int b = 0; b < n; ++b)
                      for (int a = 0; a < n; ++a)
                        {
                          double acc = 0;
                          for (int qy = 0; qy < n; ++qy)
                            acc += sh_.S[qy * n + b] * t1[(size_t(qz) * n + qy) * n + a];
                          t2[(size_t(qz) * n + b) * n + a] = acc;
                        }
                  for (int c3 = 0; c3 < nz; ++c3)
                    for (int b = 0; b < n; ++b)
                      for (int a = 0; a < n; ++a)
                        {
                          double acc = 0;
                          if (dim == 3)
                            for (int qz = 0; qz < n; ++qz)
                              acc += sh_.S[qz * n + c3] * t2[(size_t(qz) * n + b) * n + a];
                          else
                            acc = t2[size_t(b) * n + a];
                          const long gx = (long)cx * sh_.p + a, gy = (long)cy * sh_.p + b,
                                     gz = (long)cz * sh_.p + c3;
                          if (node_constrained(g, gx, gy, gz))
                            continue;
                          h[gx + g.nn[0] * (gy + g.nn[1] * gz)] += acc;
                        }
                }
            }
          }
        load_.upload(h.data(), h.size(), ctx_.stream);
      }

      Context                               ctx_;
      Shape1D                               sh_;
      std::vector<MeshLevel>                mesh_;
      std::unique_ptr<GMG<T>>               gmg_;
      std::unique_ptr<SystemMatrix<double>> fine_, rhs_, rhsv_;
      std::unique_ptr<SystemMatrix<double>> fineT_, misfit_, rhsT_; // adjoint
      std::unique_ptr<GMG<T>>               gmg_adj_;
      DBuf<double>                          dalpha_, kappa_, last_;
      std::unique_ptr<SolverFGMRES>         solver_;
      DMat                                  alpha1_, gamma1_;
      DBuf<double>                          rhs_vec_, load_, coefs_, ones_;
      DBuf<double>                          aixb_, aixg_, aixz_; // wave velocity recovery
    };
  } // namespace

  std::unique_ptr<ProblemBase> make_problem(const std::string &spec, const double *vertices,
                                            const double *kappa, int device)
  {
    const ProblemSpec ps = ProblemSpec::parse(spec);
    if (ps.mg_float)
      return std::make_unique<Problem<float>>(ps, vertices, kappa, device);
    return std::make_unique<Problem<double>>(ps, vertices, kappa, device);
  }

  template class SystemMatrix<double>;
  template class SystemMatrix<float>;
  template class PreconditionVanka<double>;
  template class PreconditionVanka<float>;
  template class MGTwoLevelTransferST<double>;
  template class MGTwoLevelTransferST<float>;
  template class MGTwoLevelBlockTransfer<double>;
  template class MGTwoLevelBlockTransfer<float>;
  template class GMG<double>;
  template class GMG<float>;
  // the solver loop for every operator/preconditioner pair of the C ABI
  template int SolverFGMRES::solve(const SystemMatrix<double> &, const GMG<double> &, double *,
                                   const double *);
  template int SolverFGMRES::solve(const SystemMatrix<double> &, const GMG<float> &, double *,
                                   const double *);
  template int SolverFGMRES::solve(const SystemMatrix<double> &, const IdentityPreconditioner &,
                                   double *, const double *);
} // namespace stfem
