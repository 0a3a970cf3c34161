// Problem specification of the space-time slab solver: what tp_01.cc:56-359
// sets up for one (refinement, degree) pair, and its parser.  Host C++ only
// (no HIP dependency).
//
// The parser accepts "key=value;key=value" strings with this library's
// snake_case keys, the camelCase keys of the reference's parameter files
// (Parameters::parse, tests/include/gmg.h:591-649) and the command-line
// options of tests/tp_01.cc:757-768 (dim, precondition_float), and a flat
// JSON object such as tests/json/tf01.json.  Unknown keys are an error.
#pragma once

#include <string>
#include <vector>

#include "status.h"
#include "time_fe.h"

namespace stfem
{
  // ------------------------------------------------------------ GMG
  struct GMGOptions
  {
    int    smoothing_steps         = 1;    // PreconditionRelaxation n_iterations
    bool   variable                = true; // MGSmootherPrecondition steps 2^(L-l)
    bool   estimate_relaxation     = true;
    int    eig_iterations          = 20;
    double eig_safety              = 1.15;
    bool   overwrite_time_transfer = false;
    bool   vanka_factored          = false;
    // On a one-cell level the Vanka patch is the whole level, so the
    // smoother is the exact inverse and s relaxation sweeps with omega equal
    // one sweep with 1 - (1 - omega)^s.
    bool   exact_single_cell       = true;
    // replay the V-cycle from a captured HIP graph: 1 yes, 0 no, -1 (default)
    // yes unless a rocprofiler-sdk tool is loaded into the process (its
    // interception of graph launches is avoided; the V-cycle then launches
    // the same kernels one by one)
    int    use_graph               = -1;
    // The V-cycle restricted to the coarse levels is a fixed linear map of
    // the restricted defect.  Below the finest level, the largest level with
    // at most this many unknowns (and everything under it) is replaced by
    // that map as one dense matrix, built at setup from the level kernels
    // and applied as one matrix-vector product per V-cycle; 0 disables.
    long   dense_coarse_dofs       = 8192;
    // relaxation sweeps on levels of at most this many cells take A u at
    // the patch nodes from the operator's cell buffer (one node gather
    // launch fewer per sweep, at the price of summing the node values in
    // the sweep kernel) where that kernel supports it; 0 disables.  The
    // default 512 covers levels of up to 8^3 cells, where the saved launch
    // latency outweighs the extra loads (on larger levels it does not)
    long   fuse_residual           = 512;
  };

  // ------------------------------------------------------------ FGMRES
  struct SolverControlData
  {
    int    maxiter = 200;
    double abstol  = 1e-12;
    double reltol  = 1e-12;
    int    basis   = 30;
    // 0: modified Gram-Schmidt as a chain of axpy+dot passes; 1: the same
    // coefficients from one pass of dots against the whole basis corrected
    // with the basis' Gram matrix (two passes over the basis per step)
    int    orthogonalization = 1;
  };

  // ------------------------------------------------------------ problem
  // Everything tp_01.cc:56-359 sets up for one (refinement, degree) pair:
  // mesh hierarchy, FP64 system and rhs matrices, MG levels, transfers, GMG,
  // FGMRES and the heat time integrator.
  struct ProblemSpec
  {
    int          dim = 2;
    int          subdiv[3] = {1, 1, 1};
    double       lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    int          refine = 2;
    int          p = -1;
    TimeStepType type = TimeStepType::DG;
    int          r = 1;
    int          nts = 1;
    double       tau = -1;
    int          space_time_mg = 1;
    int          nts_min = -1, r_min = -1;
    int          time_before_space = 0;
    int          mg_space = 1;
    int          mg_float = 0;
    int          dirichlet = 63;
    int          atomic_scatter = 0; // 1: one-pass operator with atomic adds
    int          problem = 0; // 0 heat, 1 wave (first order in time, tp_01.cc:167-197)
    int          coarse_coef = 0; // coarse-level coefficient: 0 sampled, 1 geometric, 2 harmonic, 3 arithmetic mean
    int          vanka_device_setup = 3; // factored Vanka on the GPU: 1 Jacobi, 2 divide and conquer, 4/5 Cholesky + divide and conquer/Jacobi, 6 own batched eigensolver, 3 by patch size
    int          vanka_kernel = 0; // factored sweep kernel: 0 automatic, 2 LDS-staged, 3 register-tiled
    // output components per operator cell pass (FP32 levels, FP64 operators;
    // 0: 2 for FP64 from Q4 on, where four do not fit two waves per SIMD, else 4)
    int          op_jb = 4, op_jb64 = 0;
    int          op_split = 128;         // StOpArgs::split_below of every operator
    int          op_kernel = 1;          // StOpArgs::kernel of every operator (3D cell pass)
    // time restriction: the transpose of the prolongation (default, gmg.h:557)
    // or the L2 restriction / projection matrices (gmg.h:101-113)
    int          restrict_is_transpose = 1;
    GMGOptions   mg;
    SolverControlData solver;
    int          extrapolate = 1;
    double       end_time = 1.0;
    double       frequency = 1.0;
    int          source = 0; // 0 manufactured heat source, 1 none, 2 load vector given,
                             // 3 seeded Gaussian bumps
    int          n_bumps = 4;
    long         seed    = 1;
    // per-phase timers (vmult, vanka, gmg, step) printed when the problem
    // is destroyed (the reference's printTiming)
    int          print_timing = 0;
    static ProblemSpec parse(const std::string &s);
    // every key with its resolved value, "key=value;..." (parse(to_string())
    // gives the same spec)
    std::string        to_string() const;
  };

  // The Gaussian-bump source of source=3: n bumps with centres uniform in
  // [0.2, 0.8]^3 and amplitudes uniform in [0.5, 1.5] drawn from
  // mt19937_64(seed), each uniform variate (rng() >> 11) * 2^-53 mapped
  // linearly; h(x) = sum_b a_b exp(-|x - c_b|^2 / (2 * 0.1^2)) over the
  // first dim coordinates.  bumps holds (cx, cy, cz, a) per bump.
  std::vector<double> gaussian_bumps(int n_bumps, long seed);
  double              gaussian_bumps_value(const std::vector<double> &bumps, int dim,
                                           const double *x);
} // namespace stfem
