// Exercises the library's host-only C++ (time FE algebra, spec parser, mesh
// hierarchy and patch matrices) for the AddressSanitizer / UBSan build of
// tests/test_host_sanitizers.py.  Exit status 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "mesh.h"
#include "spec.h"
#include "status.h"
#include "time_fe.h"

using namespace stfem;

static int failures = 0;
#define CHECK(c)                                                                                  \
  do                                                                                              \
    {                                                                                             \
      if (!(c))                                                                                   \
        {                                                                                         \
          std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__);                     \
          ++failures;                                                                             \
        }                                                                                         \
    }                                                                                             \
  while (0)

int main()
{
  // time FE algebra over the degrees and step counts the solver uses
  for (TimeStepType t : {TimeStepType::DG, TimeStepType::CGP})
    for (int r = (t == TimeStepType::DG ? 0 : 1); r <= 5; ++r)
      for (int nts : {1, 2, 4})
        {
          const TimeWeights w = get_fe_time_weights(t, r, 0.125, nts);
          CHECK(w.Alpha.m == n_dofs_per_interval(t, r) * nts && w.Beta.m == w.Alpha.m);
          const WaveWeights ww =
            get_fe_time_weights_wave(t, w.Alpha, w.Beta, w.Gamma, w.Zeta, 1);
          CHECK(ww.lhs_uK.m == w.Alpha.m);
          if (nts >= 2)
            {
              const DMat P = get_time_prolongation_matrix(t, r, nts);
              const DMat R = get_time_restriction_matrix(t, r, nts);
              CHECK(P.m == R.n && P.n == R.m);
            }
          if (r > (t == TimeStepType::DG ? 0 : 1))
            {
              const DMat Pk = get_time_projection_matrix(t, r - 1, r, nts);
              CHECK(Pk.m == w.Alpha.m);
            }
          for (int nsp = 1; nsp <= 4; ++nsp)
            (void)get_time_mg_sequence(nsp, r, t == TimeStepType::DG ? 0 : 1, nts,
                                       std::max(1, nts / 2), nsp % 2 == 0);
        }
  for (int n = 1; n <= 12; ++n)
    {
      double s = 0;
      for (double w : quad_gauss(n).w)
        s += w;
      CHECK(std::abs(s - 1.0) < 1e-13);
      if (n >= 2)
        (void)quad_gauss_lobatto(n);
      (void)quad_gauss_radau_right(n);
    }

  // the spec parser: valid and invalid specs, JSON, round trip
  const char *good[] = {
    "", "dim=3;refine=2;p=4;type=DG;r=3;nts=1;tau=0.25;source=3;seed=7;coarse_coef=3",
    "timeType=CGP;problemType=wave;feDegree=2;nTimestepsAtOnce=4;refinement=3;extrapolate=false",
    "{\"timeType\": \"DG\", \"problemType\": \"heat\", \"nTimestepsAtOnce\": \"2\", "
    "\"feDegree\": 1, \"refinement\": \"2\", \"doOutput\": false}"};
  for (const char *g : good)
    {
      const ProblemSpec ps = ProblemSpec::parse(g);
      CHECK(ProblemSpec::parse(ps.to_string()).to_string() == ps.to_string());
    }
  const char *bad[] = {"dim=4",          "refine=x",        "unknown=1",       "nts=3",
                       "{\"a\": [1]}",   "{\"refine\": 2",  "dim=2;p=9",       "subdiv=1,2,3,4",
                       "reltol=1e999",   "type=GCC",        "distortGrid=0.5", "key",
                       "{\"x\" 1}",      "{",               "\"",              "use_graph=5"};
  for (const char *b : bad)
    {
      bool threw = false;
      try
        {
          (void)ProblemSpec::parse(b);
        }
      catch (const Error &)
        {
          threw = true;
        }
      CHECK(threw);
    }
  CHECK(gaussian_bumps(5, 3).size() == 20);

  // mesh hierarchies with perturbed vertices, cell coefficients, every
  // coarse-coefficient mode, and the patch matrices of the Vanka setup
  for (int dim : {2, 3})
    for (int p = 1; p <= (dim == 2 ? 5 : 3); ++p)
      for (int cc = 0; cc <= 3; ++cc)
        {
          const int refine = dim == 2 ? 2 : 1, subdiv[3] = {2, 1, 1};
          const int ncx = subdiv[0] << refine, ncy = subdiv[1] << refine,
                    ncz = dim == 3 ? (subdiv[2] << refine) : 1;
          const long nv = long(ncx + 1) * (ncy + 1) * (dim == 3 ? ncz + 1 : 1);
          std::vector<double> X(size_t(nv) * dim), kappa(size_t(ncx) * ncy * ncz);
          long                id = 0;
          for (int k = 0; k <= (dim == 3 ? ncz : 0); ++k)
            for (int j = 0; j <= ncy; ++j)
              for (int i = 0; i <= ncx; ++i, ++id)
                {
                  const int    ii[3] = {i, j, k};
                  const int    nn[3] = {ncx, ncy, ncz};
                  for (int d = 0; d < dim; ++d)
                    {
                      const bool inner = ii[d] > 0 && ii[d] < nn[d];
                      X[id * dim + d] = double(ii[d]) / nn[d] +
                                        (inner ? 0.1 / nn[d] * std::sin(3.0 * id + d) : 0.0);
                    }
                }
          for (size_t c = 0; c < kappa.size(); ++c)
            kappa[c] = std::exp(2.0 * std::sin(double(c)));
          const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
          const auto   levels =
            build_mesh_levels(dim, subdiv, refine, p, X.data(), lo, hi, kappa.data(), 63 & ((1 << (2 * dim)) - 1), cc);
          CHECK(int(levels.size()) == refine + 1);
          const Shape1D       sh(p);
          std::vector<double> Kp, Mp;
          cell_patch_matrices(levels.back(), sh, Kp, Mp);
          const long nl = long(std::pow(p + 1, dim));
          CHECK(long(Kp.size()) == levels.back().g.n_cells * nl * nl);
          for (double v : Mp)
            CHECK(std::isfinite(v));
        }
  std::printf("host driver: %d failures\n", failures);
  return failures == 0 ? 0 : 1;
}
