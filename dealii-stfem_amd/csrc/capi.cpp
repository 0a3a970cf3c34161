// extern "C" boundary of libstfem_amd.so (declarations in include/stfem.h).
#include "../../include/stfem.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "capi_common.h"
#include "stfem_host.h"

using namespace stfem;

struct stfem_problem_s
{
  std::unique_ptr<ProblemBase> impl;
};

namespace
{
  void copy_mat(const DMat &m, double *out)
  {
    if (out)
      std::memcpy(out, m.a.data(), m.a.size() * sizeof(double));
  }
  TimeStepType ttype(int type)
  {
    STFEM_REQUIRE(type == STFEM_TIME_CGP || type == STFEM_TIME_DG, "unknown time step type");
    return type == STFEM_TIME_CGP ? TimeStepType::CGP : TimeStepType::DG;
  }
  ProblemBase &P(stfem_problem p)
  {
    STFEM_REQUIRE(p && p->impl, "null problem handle");
    return *p->impl;
  }
  void check_level(ProblemBase &pb, int level, bool allow_fine)
  {
    STFEM_REQUIRE((allow_fine && level == -1) || (level >= 0 && level < pb.n_levels()),
                  "level out of range");
  }
  // two HIP events, destroyed on every exit path
  struct EventPair
  {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    EventPair()
    {
      STFEM_HIP(hipEventCreate(&e0));
      if (hipEventCreate(&e1) != hipSuccess)
        {
          (void)hipEventDestroy(e0);
          throw Error(STATUS_ERR_HIP, "hipEventCreate failed");
        }
    }
    ~EventPair()
    {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
    EventPair(const EventPair &)            = delete;
    EventPair &operator=(const EventPair &) = delete;
    // milliseconds between the two records (waits for the second)
    double elapsed() const
    {
      STFEM_HIP(hipEventSynchronize(e1));
      float t = 0;
      STFEM_HIP(hipEventElapsedTime(&t, e0, e1));
      return t;
    }
  };
  // inputs may come from other streams (e.g. a framework allocator)
  void device_barrier(ProblemBase &pb)
  {
    STFEM_HIP(hipSetDevice(pb.context().device));
    STFEM_HIP(hipDeviceSynchronize());
  }
} // namespace

extern "C"
{
  int stfem_version(void) { return 100; }
  const char *stfem_last_error(void) { return last_error_message().c_str(); }

  int stfem_fe_time_weights(int type, int r, double tau, int nts, double *alpha, double *beta,
                            double *gamma, double *zeta)
  {
    return guarded([&] {
      STFEM_REQUIRE(nts >= 1, "nts must be >= 1");
      const auto w = get_fe_time_weights(ttype(type), r, tau, nts);
      copy_mat(w.Alpha, alpha);
      copy_mat(w.Beta, beta);
      copy_mat(w.Gamma, gamma);
      copy_mat(w.Zeta, zeta);
    });
  }
  int stfem_dg_weights(int r, double *mass, double *der, double *jump)
  {
    return guarded([&] {
      STFEM_REQUIRE(r >= 0, "degree must be >= 0");
      const auto w = get_dg_weights(r);
      copy_mat(w.mass, mass);
      copy_mat(w.der, der);
      copy_mat(w.jump, jump);
    });
  }
  int stfem_cg_weights(int r, double *mass, double *der)
  {
    return guarded([&] {
      STFEM_REQUIRE(r >= 1, "degree must be >= 1");
      const auto w = get_cg_weights(r);
      copy_mat(w.mass, mass);
      copy_mat(w.der, der);
    });
  }
  int stfem_fe_time_weights_wave(int type, int r, double tau, int nts, double *lhs_uK,
                                 double *lhs_uM, double *rhs_uK, double *rhs_uM, double *rhs_vM)
  {
    return guarded([&] {
      const auto w  = get_fe_time_weights(ttype(type), r, tau, 1);
      const auto ww = get_fe_time_weights_wave(ttype(type), w.Alpha, w.Beta, w.Gamma, w.Zeta, nts);
      copy_mat(ww.lhs_uK, lhs_uK);
      copy_mat(ww.lhs_uM, lhs_uM);
      copy_mat(ww.rhs_uK, rhs_uK);
      copy_mat(ww.rhs_uM, rhs_uM);
      copy_mat(ww.rhs_vM, rhs_vM);
    });
  }
  int stfem_time_prolongation(int type, int r, int nts, double *P)
  {
    return guarded([&] { copy_mat(get_time_prolongation_matrix(ttype(type), r, nts), P); });
  }
  int stfem_time_restriction(int type, int r, int nts, double *R)
  {
    return guarded([&] { copy_mat(get_time_restriction_matrix(ttype(type), r, nts), R); });
  }
  int stfem_time_projection(int type, int r_src, int r_dst, int nts, double *P, int *rows,
                            int *cols)
  {
    return guarded([&] {
      const auto m = get_time_projection_matrix(ttype(type), r_src, r_dst, nts);
      if (rows)
        *rows = m.m;
      if (cols)
        *cols = m.n;
      copy_mat(m, P);
    });
  }
  int stfem_quadrature(int kind, int n, double *points, double *weights)
  {
    return guarded([&] {
      STFEM_REQUIRE(n >= 1 && n <= 30, "quadrature size out of range");
      Quad1D q;
      if (kind == 0)
        q = quad_gauss(n);
      else if (kind == 1)
        q = quad_gauss_lobatto(n);
      else if (kind == 2)
        q = quad_gauss_radau_right(n);
      else
        throw Error(STATUS_ERR_ARG, "unknown quadrature kind");
      for (int i = 0; i < n; ++i)
        {
          if (points)
            points[i] = q.x[i];
          if (weights)
            weights[i] = q.w[i];
        }
    });
  }
  int stfem_time_mg_sequence(int n_sp_lvl, int k_max, int k_min, int nts, int nts_min,
                             int time_before_space, int *out, int *n)
  {
    return guarded([&] {
      STFEM_REQUIRE(n_sp_lvl >= 1 && nts_min >= 1 && k_max >= k_min, "bad MG sequence arguments");
      const auto s = get_time_mg_sequence(n_sp_lvl, k_max, k_min, nts, nts_min, time_before_space != 0);
      if (n)
        *n = int(s.size());
      if (out)
        for (size_t i = 0; i < s.size(); ++i)
          out[i] = int(s[i]);
    });
  }

  int stfem_spec_normalize(const char *spec, char *out, int cap)
  {
    return guarded([&] {
      const std::string n = ProblemSpec::parse(spec ? spec : "").to_string();
      STFEM_REQUIRE(out == nullptr || cap > int(n.size()), "output buffer too small");
      if (out)
        std::memcpy(out, n.c_str(), n.size() + 1);
    });
  }
  int stfem_problem_create(const char *spec, const double *vertices, const double *cell_coef,
                           int device, stfem_problem *out)
  {
    return guarded([&] {
      STFEM_REQUIRE(out, "null output handle");
      *out    = nullptr;
      auto *h = new stfem_problem_s;
      try
        {
          h->impl = make_problem(spec ? spec : "", vertices, cell_coef, device);
        }
      catch (...)
        {
          delete h;
          throw;
        }
      *out = h;
    });
  }
  int stfem_problem_destroy(stfem_problem p)
  {
    return guarded([&] { delete p; });
  }
  int stfem_problem_info(stfem_problem p, long *info)
  {
    return guarded([&] {
      auto &pb = P(p);
      long  li[12];
      pb.level_info(pb.n_levels() - 1, li);
      info[0] = pb.n_levels();
      info[1] = pb.nt;
      info[2] = pb.n_nodes();
      info[3] = li[2];
      info[4] = li[9];
      info[5] = pb.spec.nts;
    });
  }
  int stfem_problem_tau(stfem_problem p, double *tau)
  {
    return guarded([&] { *tau = P(p).tau; });
  }
  int stfem_level_info(stfem_problem p, int level, long *info)
  {
    return guarded([&] {
      auto &pb = P(p);
      check_level(pb, level, false);
      pb.level_info(level, info);
    });
  }
  int stfem_level_omega(stfem_problem p, int level, double *omega)
  {
    return guarded([&] {
      auto &pb = P(p);
      check_level(pb, level, false);
      *omega = pb.level_omega(level);
    });
  }
  int stfem_st_vmult(stfem_problem p, int level, void *dst, const void *src)
  {
    return guarded([&] {
      auto &pb = P(p);
      check_level(pb, level, true);
      device_barrier(pb);
      if (level == -1)
        pb.fine_vmult((double *)dst, (const double *)src, false);
      else
        pb.level_vmult(level, dst, src, false);
      pb.context().sync();
    });
  }
  int stfem_st_tvmult(stfem_problem p, int level, void *dst, const void *src)
  {
    return guarded([&] {
      auto &pb = P(p);
      check_level(pb, level, true);
      device_barrier(pb);
      if (level == -1)
        pb.fine_vmult((double *)dst, (const double *)src, true);
      else
        pb.level_vmult(level, dst, src, true);
      pb.context().sync();
    });
  }
  int stfem_vanka_vmult(stfem_problem p, int level, void *dst, const void *src)
  {
    return guarded([&] {
      auto &pb = P(p);
      check_level(pb, level, false);
      device_barrier(pb);
      pb.level_vanka(level, dst, src);
      pb.context().sync();
    });
  }
  int stfem_prolongate_and_add(stfem_problem p, int level, void *dst, const void *src)
  {
    return guarded([&] {
      auto &pb = P(p);
      device_barrier(pb);
      pb.level_prolongate(level, dst, src);
      pb.context().sync();
    });
  }
  int stfem_restrict_and_add(stfem_problem p, int level, void *dst, const void *src)
  {
    return guarded([&] {
      auto &pb = P(p);
      device_barrier(pb);
      pb.level_restrict(level, dst, src);
      pb.context().sync();
    });
  }
  int stfem_gmg_vmult(stfem_problem p, double *dst, const double *src)
  {
    return guarded([&] {
      auto &pb = P(p);
      device_barrier(pb);
      pb.gmg_vmult(dst, src);
      pb.context().sync();
    });
  }
  int stfem_fgmres_solve(stfem_problem p, double *x, const double *b, int *its, double *hist,
                         int hist_cap)
  {
    int rc = STFEM_OK;
    int g  = guarded([&] {
      auto &pb = P(p);
      device_barrier(pb);
      std::vector<double> h;
      rc = pb.fgmres(x, b, its, &h);
      if (hist)
        for (int i = 0; i < hist_cap && i < int(h.size()); ++i)
          hist[i] = h[i];
      if (rc == STFEM_ERR_NO_CONVERGENCE)
        last_error_message() = "FGMRES: iteration limit reached";
      else if (rc == STFEM_ERR_NOT_FINITE)
        last_error_message() = "FGMRES: residual is not finite";
    });
    return g != STFEM_OK ? g : rc;
  }
  int stfem_slab_error(stfem_problem p, const double *x, const double *prev, double time,
                       double *out)
  {
    return guarded([&] {
      auto &pb = P(p);
      device_barrier(pb);
      pb.slab_error(x, prev, time, out);
    });
  }
  int stfem_run_heat(stfem_problem p, double *stats, int *its_per_slab, int cap)
  {
    int rc = STFEM_OK;
    int g  = guarded([&] {
      auto &pb = P(p);
      device_barrier(pb);
      rc = pb.run_heat(stats, its_per_slab, cap);
      pb.context().sync();
    });
    return g != STFEM_OK ? g : rc;
  }
  int stfem_solve_slab(stfem_problem p, double *x, const double *prev, double time, int *its)
  {
    int rc = STFEM_OK;
    int g  = guarded([&] {
      auto &pb = P(p);
      device_barrier(pb);
      rc = pb.solve_slab(x, prev, time, its);
      pb.context().sync();
      if (rc == STFEM_ERR_NO_CONVERGENCE)
        last_error_message() = "FGMRES: iteration limit reached";
      else if (rc == STFEM_ERR_NOT_FINITE)
        last_error_message() = "FGMRES: residual is not finite";
    });
    return g != STFEM_OK ? g : rc;
  }
  int stfem_solve_slab_wave(stfem_problem p, double *x, double *v, const double *prev_u,
                             const double *prev_v, double time, int *its)
  {
    int rc = STFEM_OK;
    int g  = guarded([&] {
      auto &pb = P(p);
      device_barrier(pb);
      rc = pb.solve_slab_wave(x, v, prev_u, prev_v, time, its);
      pb.context().sync();
      if (rc == STFEM_ERR_NO_CONVERGENCE)
        last_error_message() = "FGMRES: iteration limit reached";
      else if (rc == STFEM_ERR_NOT_FINITE)
        last_error_message() = "FGMRES: residual is not finite";
    });
    return g != STFEM_OK ? g : rc;
  }
  int stfem_node_coordinates(stfem_problem p, double *xyz)
  {
    return guarded([&] {
      STFEM_REQUIRE(xyz, "null output");
      std::vector<double> v;
      P(p).node_coordinates_host(v);
      std::copy(v.begin(), v.end(), xyz);
    });
  }
  int stfem_forward_trajectory(stfem_problem p, int n_slabs, double *traj, int *its)
  {
    return guarded([&] {
      STFEM_REQUIRE(traj, "null trajectory");
      auto &pb = P(p);
      device_barrier(pb);
      pb.forward_trajectory(n_slabs, traj, its);
    });
  }
  int stfem_misfit(stfem_problem p, const double *obs, int n_slabs, double *misfit, int *its)
  {
    return guarded([&] {
      STFEM_REQUIRE(obs && misfit, "null argument");
      auto &pb = P(p);
      device_barrier(pb);
      *misfit = pb.misfit(obs, n_slabs, its);
    });
  }
  int stfem_adjoint_gradient(stfem_problem p, const double *obs, int n_slabs, double *grad,
                             double *misfit, int *its)
  {
    return guarded([&] {
      STFEM_REQUIRE(obs && grad && misfit, "null argument");
      auto &pb = P(p);
      device_barrier(pb);
      pb.adjoint_gradient(obs, n_slabs, grad, misfit, its);
    });
  }
  int stfem_set_cell_coefficient(stfem_problem p, const double *kappa)
  {
    return guarded([&] {
      STFEM_REQUIRE(kappa, "null coefficient");
      auto &pb = P(p);
      device_barrier(pb);
      pb.set_cell_coefficient(kappa);
    });
  }
  int stfem_set_load_vector(stfem_problem p, const double *h)
  {
    return guarded([&] {
      STFEM_REQUIRE(h, "null load vector");
      P(p).set_load_vector(h);
    });
  }
  int stfem_bytes_model(stfem_problem p, int what, int level, double *bytes)
  {
    return guarded([&] {
      auto &pb = P(p);
      if (what == 0)
        *bytes = pb.fine_vmult_bytes();
      else if (what == 1)
        {
          check_level(pb, level, false);
          *bytes = pb.vanka_sweep_bytes(level);
        }
      else
        {
          // transfer between level-1 and level: read the source, read and
          // write the destination
          STFEM_REQUIRE(what == 2 || what == 3, "unknown bytes-model selector");
          check_level(pb, level, false);
          STFEM_REQUIRE(level >= 1, "transfers need a level above 0");
          long fi[12], ci[12];
          pb.level_info(level, fi);
          pb.level_info(level - 1, ci);
          const double s = double(fi[9]), nf = double(fi[0]) * fi[1], nc = double(ci[0]) * ci[1];
          *bytes = what == 2 ? s * (nc + 2 * nf) : s * (nf + 2 * nc);
        }
    });
  }
  int stfem_blocks_to_nodes(long n_nodes, int nt, const double *blocks, double *nodes)
  {
    return guarded([&] {
      STFEM_HIP(hipDeviceSynchronize());
      launch_blocks_to_interleaved<double>(n_nodes, nt, blocks, nodes, nullptr);
      STFEM_HIP(hipDeviceSynchronize());
    });
  }
  int stfem_nodes_to_blocks(long n_nodes, int nt, const double *nodes, double *blocks)
  {
    return guarded([&] {
      STFEM_HIP(hipDeviceSynchronize());
      launch_interleaved_to_blocks<double>(n_nodes, nt, nodes, blocks, nullptr);
      STFEM_HIP(hipDeviceSynchronize());
    });
  }

  int stfem_time_kernel(stfem_problem p, int what, int level, int warmup, int reps, double *ms)
  {
    return guarded([&] {
      auto &pb = P(p);
      STFEM_REQUIRE(reps >= 1 && warmup >= 0, "bad repetition counts");
      Context &ctx = pb.context();
      STFEM_HIP(hipSetDevice(ctx.device));
      long li[12];
      if (what == 0 || what == 3)
        level = pb.n_levels() - 1;
      check_level(pb, level, false);
      STFEM_REQUIRE(what < 4 || level >= 1, "transfers need a level above 0");
      pb.level_info(level, li);
      const size_t n  = size_t(li[0]) * li[1];
      const size_t es = (what == 0 || what == 3) ? sizeof(double) : size_t(li[9]);
      // (transfers: the finer level is the larger of the two vectors)
      DBuf<unsigned char> a(n * es), b(n * es);
      {
        std::mt19937_64                  rng(12345);
        std::uniform_real_distribution<> U(-1.0, 1.0);
        if (es == sizeof(double))
          {
            std::vector<double> h(n);
            for (auto &v : h)
              v = U(rng);
            STFEM_HIP(hipMemcpy(a.get(), h.data(), n * es, hipMemcpyHostToDevice));
          }
        else
          {
            std::vector<float> h(n);
            for (auto &v : h)
              v = float(U(rng));
            STFEM_HIP(hipMemcpy(a.get(), h.data(), n * es, hipMemcpyHostToDevice));
          }
      }
      auto run = [&]() {
        switch (what)
          {
            case 0: pb.fine_vmult((double *)b.get(), (const double *)a.get(), false); break;
            case 1: pb.level_vmult(level, b.get(), a.get(), false); break;
            case 2: pb.level_vanka(level, b.get(), a.get()); break;
            case 3: pb.gmg_vmult((double *)b.get(), (const double *)a.get()); break;
            case 4: pb.level_prolongate(level, b.get(), a.get()); break;
            case 5: pb.level_restrict(level, b.get(), a.get()); break;
            case 6: pb.level_vanka_cells(level, a.get()); break;
            default: throw Error(STATUS_ERR_ARG, "unknown kernel selector");
          }
      };
      // transfers and the V-cycle accumulate into b: start from zero
      STFEM_HIP(hipMemset(b.get(), 0, n * es));
      for (int i = 0; i < warmup; ++i)
        run();
      ctx.sync();
      EventPair ev;
      STFEM_HIP(hipEventRecord(ev.e0, ctx.stream));
      for (int i = 0; i < reps; ++i)
        run();
      STFEM_HIP(hipEventRecord(ev.e1, ctx.stream));
      *ms = ev.elapsed() / reps;
    });
  }
  int stfem_phase_timings(stfem_problem p, int reset, double *ms, long *calls)
  {
    return guarded([&] {
      auto &pb = P(p);
      STFEM_REQUIRE(pb.context().timers, "phase timers are off (spec key print_timing=1)");
      STFEM_HIP(hipSetDevice(pb.context().device));
      pb.context().timers->read(ms, calls);
      if (reset)
        pb.context().timers->reset();
    });
  }
  int stfem_time_slabs(stfem_problem p, int n_slabs, double *ms, int *its)
  {
    int rc = STFEM_OK;
    int g  = guarded([&] {
      auto &pb = P(p);
      STFEM_REQUIRE(n_slabs >= 1, "n_slabs must be >= 1");
      Context &ctx = pb.context();
      STFEM_HIP(hipSetDevice(ctx.device));
      const long N  = pb.n_nodes();
      const int  nt = pb.nt;
      // slabs of one run from t = 0 to end_time; the sequence restarts from
      // a zero state after that many slabs
      const double slab_len = pb.spec.nts * pb.tau;
      const int    period   = std::max(1, int(std::lround(pb.spec.end_time / slab_len)));
      DBuf<double> x{size_t(N) * nt}, prev{size_t(N)}, sel{size_t(nt)};
      std::vector<double> s(nt, 0.0);
      s[nt - 1] = 1.0;
      sel.upload(s.data(), s.size(), ctx.stream);
      ctx.sync();
      EventPair ev;
      double    total = 0;
      int       itsum = 0;
      for (int k = 0; k < n_slabs; ++k)
        {
          if (k % period == 0)
            x.zero(ctx.stream);
          const double time = (k % period) * slab_len;
          // prev = last time value of the previous slab (tp_01.cc:637)
          launch_time_transfer<double>(N, sel.get(), 1, nt, x.get(), prev.get(), false, ctx.stream);
          STFEM_HIP(hipEventRecord(ev.e0, ctx.stream));
          int       it = 0;
          const int r  = pb.solve_slab(x.get(), prev.get(), time, &it);
          STFEM_HIP(hipEventRecord(ev.e1, ctx.stream));
          total += ev.elapsed();
          itsum += it;
          // a slab that hits the iteration limit still counts (its time and
          // iterations); the status is reported at the end
          if (r != STFEM_OK && rc == STFEM_OK)
            rc = r;
          if (r != STFEM_OK && r != STFEM_ERR_NO_CONVERGENCE)
            break;
        }
      *ms  = total;
      *its = itsum;
    });
    return g != STFEM_OK ? g : rc;
  }
}
