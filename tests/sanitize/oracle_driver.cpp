// Exercises the oracle (oracle/stfem_oracle.cpp, the tests' CPU checker)
// for the AddressSanitizer / UBSan build of tests/test_host_sanitizers.py:
// heat and wave slabs, every level operation, the adjoint and the error
// calculator on small 2D and 3D problems.  Exit status 0 when every run
// finishes with finite results.
#include <cmath>
#include <cstdio>
#include <vector>

extern "C"
{
  void  *oracle_problem_create(const char *spec, const double *coords, const double *kappa);
  void   oracle_problem_destroy(void *h);
  int    oracle_n_levels(void *h);
  void   oracle_level_info(void *h, int l, long *info);
  void   oracle_level_apply(void *h, int l, int what, double *dst, const double *src);
  void   oracle_gmg_vmult(void *h, double *dst, const double *src);
  void   oracle_fine_vmult(void *h, double *dst, const double *src, int transpose);
  void   oracle_load_vector(void *h, double time, double *out);
  int    oracle_solve_slab(void *h, double *x, const double *prev, double time, int *its);
  int    oracle_solve_slab_wave(void *h, double *x, double *v, const double *prev,
                                const double *prev_v, double time, int *its);
  int    oracle_slab_error(void *h, const double *x, const double *prev, double time,
                           double *out);
  int    oracle_adjoint_gradient(void *h, void *hT, const double *obs, int S, double *grad,
                                 double *misfit);
  int    oracle_run_heat(void *h, double *stats, int *its_per_slab, int cap);
}

static int failures = 0;
static void check(bool c, const char *what)
{
  if (!c)
    {
      std::fprintf(stderr, "check failed: %s\n", what);
      ++failures;
    }
}
static bool finite(const std::vector<double> &v)
{
  for (double x : v)
    if (!std::isfinite(x))
      return false;
  return true;
}

static void run(const char *spec, bool wave)
{
  void *h = oracle_problem_create(spec, nullptr, nullptr);
  const int L = oracle_n_levels(h);
  long      info[9];
  for (int l = 0; l < L; ++l)
    {
      oracle_level_info(h, l, info);
      const long           n = info[0] * info[1];
      std::vector<double> x(n), y(n);
      for (long i = 0; i < n; ++i)
        x[i] = std::sin(0.37 * i + l);
      for (int what = 0; what < 2; ++what)
        {
          oracle_level_apply(h, l, what, y.data(), x.data());
          check(finite(y), spec);
        }
      if (l > 0)
        {
          long ci[9];
          oracle_level_info(h, l - 1, ci);
          std::vector<double> yc(ci[0] * ci[1], 0.0), yf(n, 0.0), xc(ci[0] * ci[1], 1.0);
          oracle_level_apply(h, l, 3, yc.data(), x.data());
          oracle_level_apply(h, l, 2, yf.data(), xc.data());
          check(finite(yc) && finite(yf), spec);
        }
    }
  oracle_level_info(h, L - 1, info);
  const long          nn = info[0], nt = info[1];
  std::vector<double> x(nn * nt, 0.0), v(nn * nt, 0.0), prev(nn, 0.0), pv(nn, 0.0), y(nn * nt);
  for (long i = 0; i < nn; ++i)
    prev[i] = std::cos(0.1 * i);
  oracle_gmg_vmult(h, y.data(), x.data());
  oracle_fine_vmult(h, y.data(), x.data(), 1);
  std::vector<double> load(nn);
  oracle_load_vector(h, 0.3, load.data());
  int its = 0;
  if (wave)
    check(oracle_solve_slab_wave(h, x.data(), v.data(), prev.data(), pv.data(), 0.0, &its) == 0,
          spec);
  else
    {
      check(oracle_solve_slab(h, x.data(), prev.data(), 0.0, &its) == 0, spec);
      double err[3];
      oracle_slab_error(h, x.data(), prev.data(), 0.0, err);
      double stats[7];
      int    per[64];
      check(oracle_run_heat(h, stats, per, 64) == 0, spec);
    }
  check(finite(x) && finite(v), spec);
  oracle_problem_destroy(h);
}

int main()
{
  run("dim=2;refine=2;type=2;r=1;nts=2;r_min=0", false);
  run("dim=2;refine=2;type=1;r=2;nts=2;r_min=1;mg_float=1", false);
  run("dim=3;refine=1;type=2;r=2;nts=1;p=3;r_min=0;rhs=3;coarse_coef=3", false);
  run("dim=2;refine=2;type=1;r=2;nts=2;r_min=1;problem=1", true);
  run("dim=2;refine=2;type=2;r=1;nts=2;r_min=0;r_is_pT=0;extrapolate=0", false);
  // adjoint of two dG slabs
  {
    const char *s = "dim=2;refine=2;type=2;r=1;nts=1;r_min=0;rhs=1";
    void       *h = oracle_problem_create(s, nullptr, nullptr);
    void       *t = oracle_problem_create("dim=2;refine=2;type=2;r=1;nts=1;r_min=0;rhs=1;"
                                          "transpose_time=1",
                                          nullptr, nullptr);
    long info[9];
    oracle_level_info(h, oracle_n_levels(h) - 1, info);
    std::vector<double> obs(2 * info[0] * info[1], 0.5), grad(info[2]);
    double              J = 0;
    check(oracle_adjoint_gradient(h, t, obs.data(), 2, grad.data(), &J) == 0, "adjoint");
    check(finite(grad) && std::isfinite(J), "adjoint finite");
    oracle_problem_destroy(t);
    oracle_problem_destroy(h);
  }
  std::printf("oracle driver: %d failures\n", failures);
  return failures == 0 ? 0 : 1;
}
