// Parser of the problem specification (spec.h).
#include "spec.h"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <random>
#include <set>
#include <sstream>

namespace stfem
{
  namespace
  {
    // highest space degree with compiled kernels (MAX_N - 1)
    constexpr int max_space_degree = 5;

    std::string trim(const std::string &s)
    {
      size_t a = 0, b = s.size();
      while (a < b && std::isspace((unsigned char)s[a]))
        ++a;
      while (b > a && std::isspace((unsigned char)s[b - 1]))
        --b;
      return s.substr(a, b - a);
    }

    [[noreturn]] void bad(const std::string &msg) { throw Error(STATUS_ERR_ARG, "spec: " + msg); }

    // A flat JSON object {"key": value, ...} with string, number or boolean
    // values, as the reference's parameter files (tests/json/*.json).
    std::vector<std::pair<std::string, std::string>> parse_json(const std::string &s)
    {
      std::vector<std::pair<std::string, std::string>> out;
      size_t i = 0;
      auto   ws = [&]() {
        while (i < s.size() && std::isspace((unsigned char)s[i]))
          ++i;
      };
      auto str = [&]() {
        if (i >= s.size() || s[i] != '"')
          bad("JSON: expected a string at offset " + std::to_string(i));
        std::string v;
        for (++i; i < s.size() && s[i] != '"'; ++i)
          {
            if (s[i] == '\\')
              {
                if (++i >= s.size())
                  break;
              }
            v += s[i];
          }
        if (i >= s.size())
          bad("JSON: unterminated string");
        ++i;
        return v;
      };
      ws();
      if (i >= s.size() || s[i] != '{')
        bad("JSON: expected '{'");
      ++i;
      ws();
      if (i < s.size() && s[i] == '}')
        return out;
      while (true)
        {
          ws();
          std::string k = str();
          ws();
          if (i >= s.size() || s[i] != ':')
            bad("JSON: expected ':' after \"" + k + "\"");
          ++i;
          ws();
          std::string v;
          if (i < s.size() && s[i] == '"')
            v = str();
          else
            {
              const size_t a = i;
              while (i < s.size() && s[i] != ',' && s[i] != '}' &&
                     !std::isspace((unsigned char)s[i]))
                ++i;
              v = s.substr(a, i - a);
              if (v.empty() || v[0] == '{' || v[0] == '[')
                bad("JSON: key \"" + k + "\" needs a string, number or boolean value");
            }
          out.emplace_back(k, v);
          ws();
          if (i < s.size() && s[i] == ',')
            {
              ++i;
              continue;
            }
          if (i < s.size() && s[i] == '}')
            {
              ++i;
              ws();
              if (i != s.size())
                bad("JSON: trailing characters after the object");
              return out;
            }
          bad("JSON: expected ',' or '}' after the value of \"" + k + "\"");
        }
    }

    std::vector<std::pair<std::string, std::string>> tokenize(const std::string &s0)
    {
      const std::string s = trim(s0);
      if (!s.empty() && s[0] == '{')
        return parse_json(s);
      std::vector<std::pair<std::string, std::string>> out;
      std::stringstream ss(s);
      std::string       tok;
      while (std::getline(ss, tok, ';'))
        {
          tok = trim(tok);
          if (tok.empty())
            continue;
          const auto e = tok.find('=');
          if (e == std::string::npos)
            bad("token '" + tok + "' is not key=value");
          out.emplace_back(trim(tok.substr(0, e)), trim(tok.substr(e + 1)));
        }
      return out;
    }

    long to_long(const std::string &k, const std::string &v)
    {
      errno       = 0;
      char *end   = nullptr;
      const long x = std::strtol(v.c_str(), &end, 10);
      if (v.empty() || errno != 0 || *end != '\0')
        bad("key '" + k + "' needs an integer, got '" + v + "'");
      return x;
    }
    int to_int(const std::string &k, const std::string &v)
    {
      const long x = to_long(k, v);
      if (x < INT32_MIN || x > INT32_MAX)
        bad("key '" + k + "' out of range");
      return int(x);
    }
    double to_double(const std::string &k, const std::string &v)
    {
      errno         = 0;
      char *end     = nullptr;
      const double x = std::strtod(v.c_str(), &end);
      if (v.empty() || errno == ERANGE || *end != '\0' || !std::isfinite(x))
        bad("key '" + k + "' needs a finite number, got '" + v + "'");
      return x;
    }
    bool to_bool(const std::string &k, const std::string &v)
    {
      if (v == "true" || v == "1")
        return true;
      if (v == "false" || v == "0")
        return false;
      bad("key '" + k + "' needs true/false (or 1/0), got '" + v + "'");
    }
    template <typename F>
    void to_list(const std::string &k, const std::string &v, int max_n, F &&f)
    {
      std::stringstream s2(v);
      std::string       tok;
      int               d = 0;
      while (std::getline(s2, tok, ','))
        {
          if (d >= max_n)
            bad("key '" + k + "' has more than " + std::to_string(max_n) + " entries");
          f(d++, trim(tok));
        }
      if (d == 0)
        bad("key '" + k + "' is empty");
    }

    // camelCase keys of Parameters::parse (gmg.h:596-632) and the tp_01
    // command-line option precondition_float (tp_01.cc:767), mapped onto
    // this library's keys
    const std::map<std::string, std::string> &aliases()
    {
      static const std::map<std::string, std::string> a = {
        {"spaceTimeMg", "space_time_mg"},
        {"mgTimeBeforeSpace", "time_before_space"},
        {"timeType", "type"},
        {"problemType", "problem"},
        {"nTimestepsAtOnce", "nts"},
        {"nTimestepsAtOnceMin", "nts_min"},
        {"feDegree", "r"},
        {"feDegreeMin", "r_min"},
        {"refinement", "refine"},
        {"hyperRectLowerLeft", "lo"},
        {"hyperRectUpperRight", "hi"},
        {"subdivisions", "subdiv"},
        {"endTime", "end_time"},
        {"smoothingSteps", "smoothing_steps"},
        {"estimateRelaxation", "estimate_relaxation"},
        {"restrictIsTransposeProlongate", "restrict_is_transpose"},
        {"printTiming", "print_timing"},
        {"precondition_float", "mg_float"}};
      return a;
    }

    // keys of the reference's parameter files that do not change a single
    // slab problem: driver loops and output of tp_01.cc (doOutput,
    // functionalFile, nDegCycles, nRefCycles, sourcePoint), Chebyshev and
    // coarse-GMRES settings that the reference's solver path does not read
    // (estimate_relaxation builds its own default PreconditionerGMGAdditionalData,
    // gmg.h:797; the coarse-grid tolerances belong to the coarse GMRES of
    // gmg.h:732-768).  Their values are still checked for type.
    const std::set<std::string> &inert_keys()
    {
      static const std::set<std::string> k = {"doOutput",          "functionalFile",
                                              "nDegCycles",        "nRefCycles",
                                              "sourcePoint",       "smoothingDegree",
                                              "coarseGridMaxiter", "coarseGridAbstol",
                                              "coarseGridReltol"};
      return k;
    }
  } // namespace

  std::vector<double> gaussian_bumps(int n_bumps, long seed)
  {
    std::mt19937_64     rng{static_cast<uint64_t>(seed)};
    auto                unit = [&]() { return double(rng() >> 11) * 0x1.0p-53; };
    std::vector<double> b(4 * size_t(std::max(1, n_bumps)));
    for (size_t i = 0; i < b.size(); i += 4)
      {
        for (int d = 0; d < 3; ++d)
          b[i + d] = 0.2 + 0.6 * unit();
        b[i + 3] = 0.5 + unit();
      }
    return b;
  }

  double gaussian_bumps_value(const std::vector<double> &bumps, int dim, const double *x)
  {
    const double two_s2 = 2.0 * 0.1 * 0.1;
    double       v      = 0.0;
    for (size_t i = 0; i < bumps.size(); i += 4)
      {
        double r2 = 0.0;
        for (int d = 0; d < dim; ++d)
          r2 += (x[d] - bumps[i + d]) * (x[d] - bumps[i + d]);
        v += bumps[i + 3] * std::exp(-r2 / two_s2);
      }
    return v;
  }

  ProblemSpec ProblemSpec::parse(const std::string &s)
  {
    ProblemSpec                        ps;
    std::map<std::string, std::string> kv;
    for (const auto &[k0, v] : tokenize(s))
      {
        std::string k = k0;
        if (inert_keys().count(k))
          {
            if (k == "doOutput")
              (void)to_bool(k, v);
            else if (k == "nDegCycles" || k == "nRefCycles" || k == "smoothingDegree" ||
                     k == "coarseGridMaxiter")
              (void)to_long(k, v);
            else if (k == "coarseGridAbstol" || k == "coarseGridReltol")
              (void)to_double(k, v);
            continue;
          }
        if (k == "coarseGridSmootherType")
          {
            STFEM_UNSUPPORTED(v == "Smoother",
                              "spec: coarseGridSmootherType '" + v +
                                "' is not supported (only \"Smoother\": the coarse level is "
                                "smoothed, gmg.h:769-776)");
            continue;
          }
        if (k == "spaceTimeConvergenceTest")
          {
            STFEM_UNSUPPORTED(to_bool(k, v),
                              "spec: spaceTimeConvergenceTest=false (the region coefficient "
                              "of operators.h:280-353 with a zero source) is not supported; "
                              "pass a per-cell coefficient and source=1 instead");
            continue;
          }
        if (k == "distortGrid" || k == "distortCoeff")
          {
            STFEM_UNSUPPORTED(to_double(k, v) == 0.0,
                              "spec: " + k +
                                " != 0 is not supported (deal.II's random distortion is not "
                                "reproducible; pass perturbed vertices / a cell coefficient)");
            continue;
          }
        const auto a = aliases().find(k);
        if (a != aliases().end())
          k = a->second;
        if (kv.count(k))
          bad("key '" + k + "' given twice (also as its camelCase form)");
        kv[k] = v;
      }

    static const std::set<std::string> known = {
      "dim", "subdiv", "lo", "hi", "refine", "p", "type", "r", "nts", "tau", "space_time_mg",
      "nts_min", "r_min", "time_before_space", "mg_space", "mg_float", "dirichlet",
      "atomic_scatter", "coarse_coef", "problem", "vanka_device_setup", "vanka_kernel", "op_jb",
      "op_jb64", "op_split", "op_kernel", "restrict_is_transpose", "smoothing_steps", "variable",
      "estimate_relaxation", "eig_iters", "eig_safety", "overwrite_transfer", "vanka_factored",
      "exact_single_cell", "use_graph", "fuse_residual", "dense_coarse_dofs", "maxiter",
      "abstol", "reltol", "basis", "orthogonalization", "extrapolate", "end_time", "frequency",
      "source", "n_bumps", "seed", "print_timing"};
    for (const auto &[k, v] : kv)
      if (!known.count(k))
        bad("unknown key '" + k + "'");

    auto has  = [&](const char *k) { return kv.count(k) != 0; };
    auto geti = [&](const char *k, int &v) {
      if (has(k))
        v = to_int(k, kv[k]);
    };
    auto getl = [&](const char *k, long &v) {
      if (has(k))
        v = to_long(k, kv[k]);
    };
    auto getd = [&](const char *k, double &v) {
      if (has(k))
        v = to_double(k, kv[k]);
    };
    auto getb = [&](const char *k, bool &v) {
      if (has(k))
        v = to_bool(k, kv[k]);
    };
    auto getbi = [&](const char *k, int &v) {
      if (has(k))
        v = to_bool(k, kv[k]) ? 1 : 0;
    };
    geti("dim", ps.dim);
    if (has("subdiv"))
      to_list("subdiv", kv["subdiv"], 3,
              [&](int d, const std::string &t) { ps.subdiv[d] = to_int("subdiv", t); });
    if (has("lo"))
      to_list("lo", kv["lo"], 3, [&](int d, const std::string &t) { ps.lo[d] = to_double("lo", t); });
    if (has("hi"))
      to_list("hi", kv["hi"], 3, [&](int d, const std::string &t) { ps.hi[d] = to_double("hi", t); });
    geti("refine", ps.refine);
    geti("p", ps.p);
    if (has("type"))
      {
        const std::string &t = kv["type"];
        if (t == "CGP" || t == "1")
          ps.type = TimeStepType::CGP;
        else if (t == "DG" || t == "2")
          ps.type = TimeStepType::DG;
        else if (t == "GCC" || t == "3")
          throw Error(STATUS_ERR_UNSUPPORTED,
                      "spec: time type GCC has no time weights in the reference "
                      "(fe_time.h:292-381 handles DG and CGP only)");
        else
          bad("type must be DG or CGP, got '" + t + "'");
      }
    geti("r", ps.r);
    geti("nts", ps.nts);
    getd("tau", ps.tau);
    getbi("space_time_mg", ps.space_time_mg);
    geti("nts_min", ps.nts_min);
    geti("r_min", ps.r_min);
    getbi("time_before_space", ps.time_before_space);
    getbi("mg_space", ps.mg_space);
    getbi("mg_float", ps.mg_float);
    geti("dirichlet", ps.dirichlet);
    getbi("atomic_scatter", ps.atomic_scatter);
    geti("coarse_coef", ps.coarse_coef);
    if (has("problem"))
      {
        const std::string &t = kv["problem"];
        if (t == "heat" || t == "0")
          ps.problem = 0;
        else if (t == "wave" || t == "1")
          ps.problem = 1;
        else
          bad("problem must be heat or wave, got '" + t + "'");
      }
    geti("vanka_device_setup", ps.vanka_device_setup);
    geti("vanka_kernel", ps.vanka_kernel);
    geti("op_jb", ps.op_jb);
    geti("op_jb64", ps.op_jb64);
    geti("op_split", ps.op_split);
    geti("op_kernel", ps.op_kernel);
    getbi("restrict_is_transpose", ps.restrict_is_transpose);
    geti("smoothing_steps", ps.mg.smoothing_steps);
    getb("variable", ps.mg.variable);
    getb("estimate_relaxation", ps.mg.estimate_relaxation);
    geti("eig_iters", ps.mg.eig_iterations);
    getd("eig_safety", ps.mg.eig_safety);
    getb("overwrite_transfer", ps.mg.overwrite_time_transfer);
    getb("vanka_factored", ps.mg.vanka_factored);
    getb("exact_single_cell", ps.mg.exact_single_cell);
    geti("use_graph", ps.mg.use_graph);
    getl("fuse_residual", ps.mg.fuse_residual);
    getl("dense_coarse_dofs", ps.mg.dense_coarse_dofs);
    geti("maxiter", ps.solver.maxiter);
    getd("abstol", ps.solver.abstol);
    getd("reltol", ps.solver.reltol);
    geti("basis", ps.solver.basis);
    geti("orthogonalization", ps.solver.orthogonalization);
    getbi("extrapolate", ps.extrapolate);
    getd("end_time", ps.end_time);
    getd("frequency", ps.frequency);
    geti("source", ps.source);
    geti("n_bumps", ps.n_bumps);
    getl("seed", ps.seed);
    getbi("print_timing", ps.print_timing);

    // Parameters::parse derivations (gmg.h:637-648)
    if (ps.nts_min == -1)
      ps.nts_min = ps.nts / 2;
    ps.nts_min = std::clamp(ps.nts_min, 1, std::max(1, ps.nts));
    const int lowest = ps.type == TimeStepType::DG ? 0 : 1;
    if (ps.r_min == -1)
      ps.r_min = ps.r - 1;
    ps.r_min = std::clamp(ps.r_min, lowest, std::max(lowest, ps.r));
    // FE_Q<dim> fe(fe_degree + 1) (tp_01.cc:77)
    if (ps.p < 0)
      ps.p = ps.r + 1;

    STFEM_REQUIRE(ps.dim == 2 || ps.dim == 3, "spec: dim must be 2 or 3");
    for (int d = 0; d < ps.dim; ++d)
      {
        STFEM_REQUIRE(ps.subdiv[d] >= 1, "spec: subdivisions must be positive");
        STFEM_REQUIRE(ps.hi[d] > ps.lo[d], "spec: the box must have positive extent");
      }
    STFEM_REQUIRE(ps.refine >= 0 && ps.refine <= 12, "spec: refine must be in [0, 12]");
    STFEM_REQUIRE(ps.p >= 1 && ps.p <= max_space_degree, "spec: space degree must be in [1, 5]");
    STFEM_REQUIRE(ps.r >= (ps.type == TimeStepType::DG ? 0 : 1) && ps.r <= 8,
                  "spec: time degree out of range (DG: 0-8, CGP: 1-8)");
    STFEM_REQUIRE(ps.nts >= 1 && (ps.nts & (ps.nts - 1)) == 0,
                  "spec: n_timesteps_at_once must be a power of two");
    STFEM_REQUIRE(ps.dirichlet >= 0 && ps.dirichlet <= 63,
                  "spec: dirichlet must be a bit mask of the box faces (0-63)");
    ps.dirichlet &= (1 << (2 * ps.dim)) - 1; // faces of the box in dim dimensions
    STFEM_REQUIRE(ps.vanka_device_setup >= 0 && ps.vanka_device_setup <= 6,
                  "spec: vanka_device_setup must be in [0, 6]");
    STFEM_REQUIRE(ps.solver.orthogonalization == 0 || ps.solver.orthogonalization == 1,
                  "spec: orthogonalization must be 0 or 1");
    STFEM_REQUIRE(ps.op_jb >= 1 && ps.op_jb <= 4 && ps.op_jb64 >= 0 && ps.op_jb64 <= 4,
                  "spec: op_jb must be in [1, 4], op_jb64 in [0, 4]");
    STFEM_REQUIRE(ps.op_kernel == 0 || ps.op_kernel == 1, "spec: op_kernel must be 0 or 1");
    STFEM_REQUIRE(ps.vanka_kernel == 0 || (ps.vanka_kernel >= 2 && ps.vanka_kernel <= 5),
                  "spec: vanka_kernel must be 0, 2, 3, 4 or 5");
    STFEM_REQUIRE(ps.coarse_coef >= 0 && ps.coarse_coef <= 3, "spec: coarse_coef must be in [0, 3]");
    STFEM_REQUIRE(ps.solver.basis >= 2 && ps.solver.basis <= 60, "spec: basis must be in [2, 60]");
    STFEM_REQUIRE(ps.solver.maxiter >= 1, "spec: maxiter must be positive");
    STFEM_REQUIRE(ps.solver.abstol >= 0 && ps.solver.reltol >= 0,
                  "spec: tolerances must not be negative");
    STFEM_REQUIRE(ps.mg.smoothing_steps >= 1, "spec: smoothing_steps must be positive");
    STFEM_REQUIRE(ps.mg.eig_iterations >= 1, "spec: eig_iters must be positive");
    STFEM_REQUIRE(ps.mg.use_graph >= -1 && ps.mg.use_graph <= 1, "spec: use_graph must be -1, 0 or 1");
    STFEM_REQUIRE(ps.source >= 0 && ps.source <= 3, "spec: source must be in [0, 3]");
    STFEM_REQUIRE(ps.n_bumps >= 1 && ps.n_bumps <= 1024, "spec: n_bumps must be in [1, 1024]");
    STFEM_REQUIRE(ps.end_time > 0, "spec: end_time must be positive");
    return ps;
  }

  std::string ProblemSpec::to_string() const
  {
    std::ostringstream o;
    o.precision(17);
    o << "dim=" << dim << ";subdiv=" << subdiv[0] << ',' << subdiv[1] << ',' << subdiv[2]
      << ";lo=" << lo[0] << ',' << lo[1] << ',' << lo[2] << ";hi=" << hi[0] << ',' << hi[1] << ','
      << hi[2] << ";refine=" << refine << ";p=" << p
      << ";type=" << (type == TimeStepType::DG ? "DG" : "CGP") << ";r=" << r << ";nts=" << nts
      << ";tau=" << tau << ";space_time_mg=" << space_time_mg << ";nts_min=" << nts_min
      << ";r_min=" << r_min << ";time_before_space=" << time_before_space
      << ";mg_space=" << mg_space << ";mg_float=" << mg_float << ";dirichlet=" << dirichlet
      << ";atomic_scatter=" << atomic_scatter << ";coarse_coef=" << coarse_coef
      << ";problem=" << (problem == 1 ? "wave" : "heat")
      << ";vanka_device_setup=" << vanka_device_setup << ";vanka_kernel=" << vanka_kernel
      << ";op_jb=" << op_jb << ";op_jb64=" << op_jb64 << ";op_split=" << op_split << ";op_kernel=" << op_kernel
      << ";restrict_is_transpose=" << restrict_is_transpose
      << ";smoothing_steps=" << mg.smoothing_steps << ";variable=" << int(mg.variable)
      << ";estimate_relaxation=" << int(mg.estimate_relaxation)
      << ";eig_iters=" << mg.eig_iterations << ";eig_safety=" << mg.eig_safety
      << ";overwrite_transfer=" << int(mg.overwrite_time_transfer)
      << ";vanka_factored=" << int(mg.vanka_factored)
      << ";exact_single_cell=" << int(mg.exact_single_cell) << ";use_graph=" << mg.use_graph
      << ";fuse_residual=" << mg.fuse_residual << ";dense_coarse_dofs=" << mg.dense_coarse_dofs
      << ";maxiter=" << solver.maxiter << ";abstol=" << solver.abstol
      << ";reltol=" << solver.reltol << ";basis=" << solver.basis
      << ";orthogonalization=" << solver.orthogonalization << ";extrapolate=" << extrapolate
      << ";end_time=" << end_time << ";frequency=" << frequency << ";source=" << source
      << ";n_bumps=" << n_bumps << ";seed=" << seed << ";print_timing=" << print_timing;
    return o.str();
  }
} // namespace stfem
