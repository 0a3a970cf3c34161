// extern "C" object-level boundary of libstfem_amd.so: the reference's
// operator, smoother, transfer, multigrid and solver objects as separate
// handles that a caller composes (declarations in include/stfem.h).
#include "../../include/stfem.h"

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "capi_common.h"
#include "stfem_host.h"

using namespace stfem;

namespace
{
  DMat dmat(int m, int n, const double *a)
  {
    STFEM_REQUIRE(a, "null matrix");
    DMat M(m, n);
    std::memcpy(M.a.data(), a, sizeof(double) * size_t(m) * n);
    return M;
  }

  // default output chunking of the operator kernels (the spec keys op_jb,
  // op_jb64, op_split of stfem_problem_create at their defaults)
  template <typename T>
  void default_chunking(SystemMatrix<T> &A)
  {
    const ProblemSpec d;
    A.set_max_jb(sizeof(T) == 4 ? d.op_jb : d.op_jb64);
    A.set_split_below(d.op_split);
    A.set_op_kernel(d.op_kernel);
  }

  void check_precision(int precision)
  {
    STFEM_REQUIRE(precision == 4 || precision == 8, "precision must be 4 (float) or 8 (double)");
  }
} // namespace

// every handle keeps the context (device and stream) alive
struct stfem_context_s
{
  std::shared_ptr<Context> ctx;
};
struct stfem_space_s
{
  std::shared_ptr<Context>   ctx;
  std::shared_ptr<MeshLevel> mesh;
  std::shared_ptr<Shape1D>   sh;
};
struct stfem_stop_s
{
  std::shared_ptr<Context>               ctx;
  int                                    precision = 8;
  std::shared_ptr<SystemMatrix<float>>   f;
  std::shared_ptr<SystemMatrix<double>>  d;
};
struct stfem_vanka_s
{
  std::shared_ptr<Context>                   ctx;
  int                                        precision = 8;
  std::shared_ptr<PreconditionVanka<float>>  f;
  std::shared_ptr<PreconditionVanka<double>> d;
};
struct stfem_transfer_s
{
  std::shared_ptr<Context> ctx;
  int                      precision = 8;
  TwoLevelTransfer<float>  f;
  TwoLevelTransfer<double> d;
};
struct stfem_gmg_s
{
  std::shared_ptr<Context>     ctx;
  int                          precision = 8;
  std::shared_ptr<GMG<float>>  f;
  std::shared_ptr<GMG<double>> d;
};
struct stfem_fgmres_s
{
  std::shared_ptr<Context>      ctx;
  long                          n = 0;
  std::shared_ptr<SolverFGMRES> s;
  std::vector<double>           last_hist;
};

namespace
{
  Context &C(const std::shared_ptr<Context> &c)
  {
    STFEM_REQUIRE(c, "null context");
    STFEM_HIP(hipSetDevice(c->device));
    return *c;
  }
  // inputs may come from other streams: wait for all prior device work
  void barrier(Context &c) { STFEM_HIP(hipDeviceSynchronize()); (void)c; }

  template <typename T>
  std::shared_ptr<GMG<T>> make_gmg(Context &ctx, int n, const stfem_stop *ops,
                                   const stfem_vanka *sm, const stfem_transfer *tr,
                                   const GMGOptions &opt, const std::vector<double> *omega)
  {
    std::vector<GMGLevel<T>>         lv(n);
    std::vector<TwoLevelTransfer<T>> tt(n);
    for (int l = 0; l < n; ++l)
      {
        STFEM_REQUIRE(ops[l] && sm[l], "gmg: null level operator or smoother");
        STFEM_REQUIRE(ops[l]->precision == int(sizeof(T)) && sm[l]->precision == int(sizeof(T)),
                      "gmg: all levels need the same precision");
        if constexpr (std::is_same_v<T, float>)
          {
            lv[l].A = ops[l]->f;
            lv[l].V = sm[l]->f;
          }
        else
          {
            lv[l].A = ops[l]->d;
            lv[l].V = sm[l]->d;
          }
        STFEM_REQUIRE(lv[l].A->n_nodes() == lv[l].V->n_nodes() &&
                        lv[l].A->n_blocks() == lv[l].V->n_blocks(),
                      "gmg: smoother and operator of a level differ in size");
        lv[l].r            = -1;
        lv[l].nts          = 1;
        lv[l].type_to_here = TimeMGType::none;
        if (l > 0)
          {
            STFEM_REQUIRE(tr[l] && tr[l]->precision == int(sizeof(T)),
                          "gmg: null transfer or wrong precision");
            if constexpr (std::is_same_v<T, float>)
              tt[l] = tr[l]->f;
            else
              tt[l] = tr[l]->d;
          }
      }
    return std::make_shared<GMG<T>>(ctx, std::move(lv), std::move(tt), opt, omega);
  }
} // namespace

extern "C"
{
  // ------------------------------------------------------------ context
  int stfem_context_create(int device, stfem_context *out)
  {
    return guarded([&] {
      STFEM_REQUIRE(out, "null output handle");
      *out   = nullptr;
      auto h = std::make_unique<stfem_context_s>();
      h->ctx = std::make_shared<Context>(device);
      *out   = h.release();
    });
  }
  int stfem_context_destroy(stfem_context c)
  {
    return guarded([&] { delete c; });
  }

  // ------------------------------------------------------------ spaces
  int stfem_space_hierarchy_create(stfem_context c, const char *spec, const double *vertices,
                                   const double *cell_coef, int cap, stfem_space *spaces, int *n)
  {
    return guarded([&] {
      STFEM_REQUIRE(c && spaces && n, "null argument");
      Context          &ctx = C(c->ctx);
      const ProblemSpec ps  = ProblemSpec::parse(spec ? spec : "");
      auto levels = build_mesh_levels(ps.dim, ps.subdiv, ps.refine, ps.p, vertices, ps.lo, ps.hi,
                                      cell_coef, ps.dirichlet, ps.coarse_coef);
      STFEM_REQUIRE(cap >= int(levels.size()), "space hierarchy: output array too small");
      (void)ctx;
      auto sh = std::make_shared<Shape1D>(ps.p);
      for (size_t s = 0; s < levels.size(); ++s)
        {
          auto *h = new stfem_space_s;
          h->ctx  = c->ctx;
          h->mesh = std::make_shared<MeshLevel>(std::move(levels[s]));
          h->sh   = sh;
          spaces[s] = h;
        }
      *n = int(levels.size());
    });
  }
  int stfem_space_destroy(stfem_space s)
  {
    return guarded([&] { delete s; });
  }
  int stfem_space_info(stfem_space s, long *info)
  {
    return guarded([&] {
      STFEM_REQUIRE(s && info, "null argument");
      const GridDesc &g = s->mesh->g;
      info[0]           = g.dim;
      info[1]           = g.p;
      info[2]           = g.nc[0];
      info[3]           = g.nc[1];
      info[4]           = g.nc[2];
      info[5]           = g.n_nodes;
      info[6]           = g.n_cells;
      info[7]           = g.dirichlet;
    });
  }
  int stfem_space_constrained(stfem_space s, unsigned char *mask)
  {
    return guarded([&] {
      STFEM_REQUIRE(s && mask, "null argument");
      const GridDesc &g = s->mesh->g;
      for (long k = 0; k < g.nn[2]; ++k)
        for (long j = 0; j < g.nn[1]; ++j)
          for (long i = 0; i < g.nn[0]; ++i)
            mask[i + g.nn[0] * (j + g.nn[1] * k)] = node_constrained(g, i, j, k) ? 1 : 0;
    });
  }

  // ------------------------------------------------------------ SystemMatrix
  int stfem_stop_create(stfem_space s, int nt_rows, int nt_cols, const double *alpha,
                        const double *beta, int precision, stfem_stop *out)
  {
    return guarded([&] {
      STFEM_REQUIRE(s && out, "null argument");
      STFEM_REQUIRE(nt_rows >= 1 && nt_cols >= 1 && nt_rows <= 16 && nt_cols <= 16,
                    "time blocks must be in [1, 16]");
      check_precision(precision);
      *out         = nullptr;
      Context &ctx = C(s->ctx);
      const DMat A = dmat(nt_rows, nt_cols, alpha), B = dmat(nt_rows, nt_cols, beta);
      auto       h = std::make_unique<stfem_stop_s>();
      h->ctx       = s->ctx;
      h->precision = precision;
      if (precision == 4)
        {
          h->f = std::make_shared<SystemMatrix<float>>(ctx, *s->mesh, A, B);
          default_chunking(*h->f);
        }
      else
        {
          h->d = std::make_shared<SystemMatrix<double>>(ctx, *s->mesh, A, B);
          default_chunking(*h->d);
        }
      ctx.sync();
      *out = h.release();
    });
  }
  int stfem_stop_destroy(stfem_stop op)
  {
    return guarded([&] { delete op; });
  }
  int stfem_stop_info(stfem_stop op, long *info)
  {
    return guarded([&] {
      STFEM_REQUIRE(op && info, "null argument");
      const bool f = op->precision == 4;
      info[0]      = f ? op->f->n_nodes() : op->d->n_nodes();
      info[1]      = f ? op->f->n_blocks() : op->d->n_blocks();
      info[2]      = f ? op->f->alpha().n : op->d->alpha().n;
      info[3]      = op->precision;
    });
  }
  int stfem_stop_vmult(stfem_stop op, void *dst, const void *src)
  {
    return guarded([&] {
      STFEM_REQUIRE(op && dst && src, "null argument");
      Context &ctx = C(op->ctx);
      barrier(ctx);
      if (op->precision == 4)
        op->f->vmult((float *)dst, (const float *)src);
      else
        op->d->vmult((double *)dst, (const double *)src);
      ctx.sync();
    });
  }
  int stfem_stop_vmult_noalloc(stfem_stop op, void *dst, const void *src)
  {
    return guarded([&] {
      STFEM_REQUIRE(op && dst && src, "null argument");
      Context &ctx = C(op->ctx);
      barrier(ctx);
      if (op->precision == 4)
        vmult_noalloc(*op->f, ctx.device, (float *)dst, (const float *)src);
      else
        vmult_noalloc(*op->d, ctx.device, (double *)dst, (const double *)src);
      ctx.sync();
    });
  }
  int stfem_stop_tvmult(stfem_stop op, void *dst, const void *src)
  {
    return guarded([&] {
      STFEM_REQUIRE(op && dst && src, "null argument");
      Context &ctx = C(op->ctx);
      barrier(ctx);
      if (op->precision == 4)
        op->f->Tvmult((float *)dst, (const float *)src);
      else
        op->d->Tvmult((double *)dst, (const double *)src);
      ctx.sync();
    });
  }
  int stfem_stop_vmult_add(stfem_stop op, void *dst, const void *src)
  {
    return guarded([&] {
      STFEM_REQUIRE(op && dst && src, "null argument");
      Context &ctx = C(op->ctx);
      barrier(ctx);
      if (op->precision == 4)
        op->f->vmult_add_nx1((float *)dst, (const float *)src);
      else
        op->d->vmult_add_nx1((double *)dst, (const double *)src);
      ctx.sync();
    });
  }

  // ------------------------------------------------------------ Vanka
  int stfem_vanka_create(stfem_space s, int nt, const double *alpha, const double *beta,
                         int precision, int form, int device_setup, stfem_vanka *out)
  {
    return guarded([&] {
      STFEM_REQUIRE(s && out, "null argument");
      STFEM_REQUIRE(nt >= 1 && nt <= 16, "time blocks must be in [1, 16]");
      STFEM_REQUIRE(form == 0 || form == 1, "form must be 0 (explicit) or 1 (factored)");
      STFEM_REQUIRE(device_setup >= 0 && device_setup <= 5, "device_setup must be in [0, 5]");
      check_precision(precision);
      *out         = nullptr;
      Context &ctx = C(s->ctx);
      const DMat A = dmat(nt, nt, alpha), B = dmat(nt, nt, beta);
      auto       h = std::make_unique<stfem_vanka_s>();
      h->ctx       = s->ctx;
      h->precision = precision;
      if (precision == 4)
        h->f = std::make_shared<PreconditionVanka<float>>(
          ctx, *s->mesh, *s->sh, A, B,
          form ? PreconditionVanka<float>::Form::factored :
                 PreconditionVanka<float>::Form::explicit_inverse,
          device_setup);
      else
        h->d = std::make_shared<PreconditionVanka<double>>(
          ctx, *s->mesh, *s->sh, A, B,
          form ? PreconditionVanka<double>::Form::factored :
                 PreconditionVanka<double>::Form::explicit_inverse,
          device_setup);
      ctx.sync();
      *out = h.release();
    });
  }
  int stfem_vanka_destroy(stfem_vanka v)
  {
    return guarded([&] { delete v; });
  }
  int stfem_vanka_apply(stfem_vanka v, void *dst, const void *src)
  {
    return guarded([&] {
      STFEM_REQUIRE(v && dst && src, "null argument");
      Context &ctx = C(v->ctx);
      barrier(ctx);
      if (v->precision == 4)
        v->f->vmult((float *)dst, (const float *)src);
      else
        v->d->vmult((double *)dst, (const double *)src);
      ctx.sync();
    });
  }

  // ------------------------------------------------------------ transfers
  int stfem_time_transfer_create(stfem_space s, int type, int r, int nts, int kind,
                                 int restrict_is_transpose, int precision, stfem_transfer *out)
  {
    return guarded([&] {
      STFEM_REQUIRE(s && out, "null argument");
      STFEM_REQUIRE(type == STFEM_TIME_CGP || type == STFEM_TIME_DG, "unknown time step type");
      STFEM_REQUIRE(kind == 1 || kind == 2, "kind must be 1 (tau) or 2 (k)");
      STFEM_REQUIRE(nts >= 1 && (kind == 2 || nts >= 2), "tau transfers need nts >= 2");
      check_precision(precision);
      *out                 = nullptr;
      Context           &ctx = C(s->ctx);
      const TimeStepType tt  = type == STFEM_TIME_CGP ? TimeStepType::CGP : TimeStepType::DG;
      const TimeMGType   k   = kind == 1 ? TimeMGType::tau : TimeMGType::k;
      const long         nn  = s->mesh->g.n_nodes;
      auto               h   = std::make_unique<stfem_transfer_s>();
      h->ctx                 = s->ctx;
      h->precision           = precision;
      if (precision == 4)
        h->f.time = std::make_shared<MGTwoLevelTransferST<float>>(ctx, tt, r, nts, k, nn, false,
                                                                  restrict_is_transpose != 0);
      else
        h->d.time = std::make_shared<MGTwoLevelTransferST<double>>(ctx, tt, r, nts, k, nn, false,
                                                                   restrict_is_transpose != 0);
      ctx.sync();
      *out = h.release();
    });
  }
  int stfem_space_transfer_create(stfem_space fine, stfem_space coarse, int nt, int precision,
                                  stfem_transfer *out)
  {
    return guarded([&] {
      STFEM_REQUIRE(fine && coarse && out, "null argument");
      STFEM_REQUIRE(nt >= 1 && nt <= 16, "time blocks must be in [1, 16]");
      const GridDesc &f = fine->mesh->g, &c = coarse->mesh->g;
      STFEM_REQUIRE(f.dim == c.dim && f.p == c.p && f.dirichlet == c.dirichlet,
                    "space transfer: the spaces differ in dimension, degree or boundary");
      for (int d = 0; d < f.dim; ++d)
        STFEM_REQUIRE(f.nc[d] == 2 * c.nc[d],
                      "space transfer: the fine space must refine the coarse one once");
      check_precision(precision);
      *out         = nullptr;
      Context &ctx = C(fine->ctx);
      auto     h   = std::make_unique<stfem_transfer_s>();
      h->ctx       = fine->ctx;
      h->precision = precision;
      if (precision == 4)
        h->f.space = std::make_shared<MGTwoLevelBlockTransfer<float>>(ctx, f, c, *fine->sh, nt);
      else
        h->d.space = std::make_shared<MGTwoLevelBlockTransfer<double>>(ctx, f, c, *fine->sh, nt);
      ctx.sync();
      *out = h.release();
    });
  }
  int stfem_transfer_destroy(stfem_transfer t)
  {
    return guarded([&] { delete t; });
  }
  int stfem_transfer_prolongate_and_add(stfem_transfer t, void *dst, const void *src)
  {
    return guarded([&] {
      STFEM_REQUIRE(t && dst && src, "null argument");
      Context &ctx = C(t->ctx);
      barrier(ctx);
      if (t->precision == 4)
        t->f.prolongate_and_add((float *)dst, (const float *)src);
      else
        t->d.prolongate_and_add((double *)dst, (const double *)src);
      ctx.sync();
    });
  }
  int stfem_transfer_restrict_and_add(stfem_transfer t, void *dst, const void *src)
  {
    return guarded([&] {
      STFEM_REQUIRE(t && dst && src, "null argument");
      Context &ctx = C(t->ctx);
      barrier(ctx);
      if (t->precision == 4)
        t->f.restrict_and_add((float *)dst, (const float *)src);
      else
        t->d.restrict_and_add((double *)dst, (const double *)src);
      ctx.sync();
    });
  }

  // ------------------------------------------------------------ GMG
  int stfem_gmg_create(int n_levels, const stfem_stop *ops, const stfem_vanka *smoothers,
                       const stfem_transfer *transfers, const char *options,
                       const double *omega, stfem_gmg *out)
  {
    return guarded([&] {
      STFEM_REQUIRE(n_levels >= 1 && ops && smoothers && out, "null argument");
      STFEM_REQUIRE(n_levels == 1 || transfers, "gmg: transfers needed between levels");
      STFEM_REQUIRE(ops[0], "gmg: null level operator");
      *out = nullptr;
      // GMG options from the spec keys smoothing_steps, variable,
      // estimate_relaxation, eig_iters, eig_safety, exact_single_cell,
      // use_graph, dense_coarse_dofs, fuse_residual (others are rejected)
      const std::string o = options ? options : "";
      const ProblemSpec ps = ProblemSpec::parse(o);
      {
        const std::string def = ProblemSpec::parse("").to_string(), got = ps.to_string();
        // only the multigrid keys may differ from the defaults
        const char *mg_keys[] = {"smoothing_steps", "variable",          "estimate_relaxation",
                                 "eig_iters",       "eig_safety",        "exact_single_cell",
                                 "use_graph",       "dense_coarse_dofs", "fuse_residual"};
        auto strip = [&](std::string s) {
          std::string r;
          size_t      a = 0;
          while (a < s.size())
            {
              size_t b = s.find(';', a);
              if (b == std::string::npos)
                b = s.size();
              const std::string kv = s.substr(a, b - a);
              bool              mg = false;
              for (const char *k : mg_keys)
                mg = mg || kv.rfind(std::string(k) + "=", 0) == 0;
              if (!mg)
                r += kv + ";";
              a = b + 1;
            }
          return r;
        };
        STFEM_REQUIRE(strip(def) == strip(got),
                      "gmg: options may only set smoothing_steps, variable, "
                      "estimate_relaxation, eig_iters, eig_safety, exact_single_cell, "
                      "use_graph, dense_coarse_dofs, fuse_residual");
      }
      std::vector<double> om;
      if (omega)
        om.assign(omega, omega + n_levels);
      auto    h    = std::make_unique<stfem_gmg_s>();
      h->ctx       = ops[0]->ctx;
      h->precision = ops[0]->precision;
      Context &ctx = C(h->ctx);
      for (int l = 0; l < n_levels; ++l)
        STFEM_REQUIRE(ops[l] && ops[l]->ctx == h->ctx && smoothers[l] &&
                        smoothers[l]->ctx == h->ctx &&
                        (l == 0 || (transfers[l] && transfers[l]->ctx == h->ctx)),
                      "gmg: every object must belong to the same context");
      if (h->precision == 4)
        h->f = make_gmg<float>(ctx, n_levels, ops, smoothers, transfers, ps.mg,
                               omega ? &om : nullptr);
      else
        h->d = make_gmg<double>(ctx, n_levels, ops, smoothers, transfers, ps.mg,
                                omega ? &om : nullptr);
      ctx.sync();
      *out = h.release();
    });
  }
  int stfem_gmg_destroy(stfem_gmg g)
  {
    return guarded([&] { delete g; });
  }
  int stfem_gmg_omega(stfem_gmg g, int level, double *omega)
  {
    return guarded([&] {
      STFEM_REQUIRE(g && omega, "null argument");
      const int n = g->precision == 4 ? g->f->n_levels() : g->d->n_levels();
      STFEM_REQUIRE(level >= 0 && level < n, "level out of range");
      *omega = g->precision == 4 ? g->f->level(level).omega : g->d->level(level).omega;
    });
  }
  int stfem_gmg_apply(stfem_gmg g, double *dst, const double *src)
  {
    return guarded([&] {
      STFEM_REQUIRE(g && dst && src, "null argument");
      Context &ctx = C(g->ctx);
      barrier(ctx);
      if (g->precision == 4)
        g->f->vmult(dst, src);
      else
        g->d->vmult(dst, src);
      ctx.sync();
    });
  }

  // ------------------------------------------------------------ FGMRES
  int stfem_fgmres_create(stfem_context c, long n, int maxiter, double abstol, double reltol,
                          int basis, int orthogonalization, stfem_fgmres *out)
  {
    return guarded([&] {
      STFEM_REQUIRE(c && out && n >= 1, "null argument or empty system");
      STFEM_REQUIRE(maxiter >= 1 && basis >= 2 && basis <= 60 && abstol >= 0 && reltol >= 0 &&
                      (orthogonalization == 0 || orthogonalization == 1),
                    "fgmres: maxiter >= 1, basis in [2, 60], tolerances >= 0, "
                    "orthogonalization 0 or 1");
      *out         = nullptr;
      Context &ctx = C(c->ctx);
      SolverControlData sc;
      sc.maxiter           = maxiter;
      sc.abstol            = abstol;
      sc.reltol            = reltol;
      sc.basis             = basis;
      sc.orthogonalization = orthogonalization;
      auto h               = std::make_unique<stfem_fgmres_s>();
      h->ctx               = c->ctx;
      h->n                 = n;
      h->s                 = std::make_shared<SolverFGMRES>(ctx, n, sc);
      ctx.sync();
      *out = h.release();
    });
  }
  int stfem_fgmres_destroy(stfem_fgmres s)
  {
    return guarded([&] { delete s; });
  }
  int stfem_fgmres_run(stfem_fgmres s, stfem_stop A, stfem_gmg P, double *x, const double *b,
                       int *its, double *hist, int hist_cap)
  {
    int rc = STFEM_OK;
    int g  = guarded([&] {
      STFEM_REQUIRE(s && A && x && b, "null argument");
      STFEM_REQUIRE(A->precision == 8, "fgmres: the system operator must be FP64");
      STFEM_REQUIRE(A->d->m() == s->n, "fgmres: operator and solver sizes differ");
      STFEM_REQUIRE(!P || (P->ctx == s->ctx && A->ctx == s->ctx),
                    "fgmres: objects of different contexts");
      Context &ctx = C(s->ctx);
      barrier(ctx);
      if (P && P->precision == 4)
        rc = s->s->solve(*A->d, *P->f, x, b);
      else if (P)
        rc = s->s->solve(*A->d, *P->d, x, b);
      else
        rc = s->s->solve(*A->d, IdentityPreconditioner{&ctx, s->n}, x, b);
      ctx.sync();
      if (its)
        *its = s->s->last_step();
      if (hist)
        {
          const auto &h = s->s->history();
          for (int i = 0; i < hist_cap && i < int(h.size()); ++i)
            hist[i] = h[i];
        }
    });
    if (g != STFEM_OK)
      return g;
    if (rc == STATUS_ERR_NOT_FINITE)
      last_error_message() = "fgmres: residual not finite";
    else if (rc == STATUS_ERR_NO_CONVERGE)
      last_error_message() = "fgmres: iteration limit reached";
    return rc;
  }
}

