/* C ABI of the MI355X space-time multigrid slab solver (libstfem_amd.so).
 *
 * Drop-in boundary for the hot path of nlsmrg/dealii-stfem: the space-time
 * operator, the Vanka smoother, the space/time multigrid transfers, the GMG
 * V-cycle and the FGMRES slab solve.  Each entry point names the reference
 * interface it replaces (file:line in nlsmrg/dealii-stfem).
 *
 * Conventions
 *   - Status: 0 on success, negative STFEM_ERR_* on failure; the message
 *     of the last failure on the calling thread is stfem_last_error().
 *   - Vectors passed as void* / double* to operator entry points are DEVICE
 *     pointers on the problem's device, in node-major layout: the nt time
 *     values of a node are contiguous, x[node * nt + t].  Nodes are the
 *     Gauss-Lobatto points of FE_Q(p) on the structured mesh, numbered
 *     lexicographically (x fastest).  stfem_blocks_to_nodes converts the
 *     reference's BlockVector layout (block t holds all nodes).
 *   - MG level vectors use the level precision (float when the spec sets
 *     mg_float=1, double otherwise); level -1 denotes the FP64 fine system.
 *   - Calls are synchronous with respect to the host: inputs are read after
 *     all prior device work completes and outputs are complete on return.
 *   - A problem handle must not be used from two threads at once (the
 *     reference is single-threaded per MPI rank).
 */
#ifndef STFEM_H
#define STFEM_H

#ifdef __cplusplus
extern "C" {
#endif

#define STFEM_OK 0
#define STFEM_ERR_ARG -1
#define STFEM_ERR_HIP -2
#define STFEM_ERR_NO_CONVERGENCE -3
#define STFEM_ERR_UNSUPPORTED -4
#define STFEM_ERR_NOT_FINITE -5
#define STFEM_ERR_INTERNAL -6

#define STFEM_TIME_CGP 1
#define STFEM_TIME_DG 2

typedef struct stfem_problem_s *stfem_problem;

int         stfem_version(void);
const char *stfem_last_error(void);

/* ---- time finite element algebra (host, row-major outputs) ---------- */
/* get_fe_time_weights, tests/include/fe_time.h:292-349.  Sizes: alpha and
 * beta nt x nt, gamma and zeta nt x 1, nt = (r+1)*nts (DG) or r*nts (CGP). */
int stfem_fe_time_weights(int type, int r, double tau, int nts, double *alpha, double *beta,
                          double *gamma, double *zeta);
/* get_dg_weights / get_cg_weights, fe_time.h:473-555 */
int stfem_dg_weights(int r, double *mass, double *der, double *jump);
int stfem_cg_weights(int r, double *mass, double *der);
/* get_fe_time_weights_wave (single-step weights of degree r, tau scaled),
 * fe_time.h:98-246 */
int stfem_fe_time_weights_wave(int type, int r, double tau, int nts, double *lhs_uK,
                               double *lhs_uM, double *rhs_uK, double *rhs_uM, double *rhs_vM);
/* get_time_prolongation_matrix / get_time_restriction_matrix,
 * fe_time.h:625-716: P is (ni*nts) x (ni*nts/2), R its shape transposed */
int stfem_time_prolongation(int type, int r, int nts, double *P);
int stfem_time_restriction(int type, int r, int nts, double *R);
/* get_time_projection_matrix, fe_time.h:567-623; P may be NULL to query
 * the shape */
int stfem_time_projection(int type, int r_src, int r_dst, int nts, double *P, int *rows,
                          int *cols);
/* kind 0 Gauss, 1 Gauss-Lobatto, 2 Gauss-Radau (right end point) on [0,1] */
int stfem_quadrature(int kind, int n, double *points, double *weights);
/* get_time_mg_sequence, fe_time.h:43-69 (1 tau, 2 k, 3 none); out may be
 * NULL to query the length */
int stfem_time_mg_sequence(int n_sp_lvl, int k_max, int k_min, int nts, int nts_min,
                           int time_before_space, int *out, int *n);

/* ---- problem: hierarchy + solver ------------------------------------ */
/* Builds what tests/tp_01.cc:56-359 builds for one (refinement, degree):
 * mesh hierarchy of a subdivided box refined `refine` times, FP64 system
 * and rhs matrices (SystemMatrix, operators.h:71-256), MG levels with
 * PreconditionVanka (gmg.h:402-540), transfers (gmg.h:35-400), GMG
 * (gmg.h:652-906) and FGMRES (time_integrators.h:44, :113).
 * spec: "key=value;..." (or a JSON parameter file, see stfem_spec_normalize)
 * with keys dim, subdiv (a,b,c), lo, hi, refine, p,
 * type (DG|CGP), r, nts, tau, nts_min, r_min, space_time_mg,
 * time_before_space, mg_space, mg_float, smoothing_steps, variable,
 * estimate_relaxation, eig_iters, eig_safety, overwrite_transfer,
 * vanka_factored, exact_single_cell, use_graph (1 replay the V-cycle as a
 * HIP graph, 0 launch it kernel by kernel, -1 (default) the graph unless a
 * rocprofiler-sdk tool is loaded), maxiter, abstol, reltol,
 * basis, extrapolate (1: every block of the start vector is the previous
 * slab's last value, 0: zero; time_integrators.h:96-104), end_time,
 * frequency, restrict_is_transpose (1 default: time restriction = P^T; 0:
 * the L2 restriction / projection matrices, gmg.h:101-113), print_timing,
 * source (0 manufactured heat
 * source, 1 none, 2 load vector set by stfem_set_load_vector, 3 seeded
 * Gaussian bumps with n_bumps, seed), dirichlet (face bit mask), problem
 * (heat, or wave: the first-order-in-time wave system of tp_01.cc:167-197,
 * solved with stfem_solve_slab_wave), atomic_scatter, vanka_device_setup
 * (0 host, 1 device with the batched Jacobi generalized eigensolver, 2
 * device with divide and conquer, 4/5 device with a Cholesky reduction to a
 * standard problem and divide and conquer/Jacobi, 3 (default) mode 4 for
 * patches above 64 nodes, 5 otherwise), dense_coarse_dofs, coarse_coef (0 the
 * coarse-level coefficient sampled at the quadrature points, 1/2/3 its
 * geometric/harmonic/arithmetic mean over each point's share of the cell),
 * vanka_kernel (0 automatic, 2 LDS-staged, 3 register-tiled, 4 pipelined),
 * fuse_residual (relaxation sweeps on levels of at most this many cells sum
 * A u at the patch nodes from the operator's cell buffer instead of a node
 * gather; default 512, 0 disables), orthogonalization (1, default: the
 * modified Gram-Schmidt coefficients of
 * an FGMRES step from one pass of dots against the basis, corrected with
 * the basis' Gram matrix; 0: the chain of axpy+dot passes),
 * op_kernel (3D cell pass with at most 4 input time components: 1, the
 * default, cells on groups of wavefronts, all output components in one
 * pass; 0 the workgroup kernel, which 2D operators always use),
 * op_jb/op_jb64 (output components per pass of the workgroup kernel, FP32
 * levels/FP64 operators; op_jb64=0, the default, picks 2 from Q4 on and 4
 * below), op_split (operators with fewer cell workgroups than this, 128 by
 * default, compute one output time component per workgroup).
 * Unknown keys are an error (STFEM_ERR_ARG).
 * vertices: finest-level vertex coordinates ((nc+1)^dim points, dim each,
 * x fastest) or NULL for the uniform box; cell_coef: finest-level per-cell
 * coefficient or NULL for 1. */
int stfem_problem_create(const char *spec, const double *vertices, const double *cell_coef,
                         int device, stfem_problem *out);
/* Parameters::parse (tests/include/gmg.h:591-649) without building anything:
 * the spec of stfem_problem_create with every key resolved, as
 * "key=value;...", into out (cap bytes including the terminating zero; out
 * NULL only checks the spec).  Accepted besides this library's keys: the
 * camelCase keys of the reference's parameter files (spaceTimeMg,
 * mgTimeBeforeSpace, timeType, problemType, nTimestepsAtOnce[Min],
 * feDegree[Min] (the time degree; the space degree defaults to feDegree+1 as
 * in tp_01.cc:77), refinement, extrapolate, hyperRectLowerLeft/UpperRight,
 * subdivisions, endTime, smoothingSteps, estimateRelaxation,
 * restrictIsTransposeProlongate, variable, printTiming), tp_01's
 * precondition_float, and a whole JSON parameter file (a spec starting with
 * '{').  Driver-only keys (doOutput, functionalFile, nDegCycles, nRefCycles,
 * sourcePoint, smoothingDegree, coarseGrid{Maxiter,Abstol,Reltol}) are
 * checked and have no effect.  Unknown keys, malformed values and keys
 * given twice return STFEM_ERR_ARG; coarseGridSmootherType other than
 * "Smoother", spaceTimeConvergenceTest=false, nonzero distortGrid /
 * distortCoeff and timeType GCC return STFEM_ERR_UNSUPPORTED. */
int stfem_spec_normalize(const char *spec, char *out, int cap);
int stfem_problem_destroy(stfem_problem p);
/* info[0..5] = n_levels, nt, n_nodes(fine), n_cells(fine), sizeof(level
 * number), n_timesteps_at_once */
int stfem_problem_info(stfem_problem p, long *info);
int stfem_problem_tau(stfem_problem p, double *tau);
/* info[0..11] = n_nodes, nt, n_cells, r, nts, mg type to this level,
 * cells x, y, z, sizeof(number), 1 if the level lies under the dense coarse
 * operator (spec key dense_coarse_dofs: the V-cycle from that level down is
 * one precomputed matrix) else 0, and on the level where that matrix is
 * applied its order (else 0) */
int stfem_level_info(stfem_problem p, int level, long *info);
/* relaxation parameter of the level (GMG::estimate_relaxation, gmg.h:792-844) */
int stfem_level_omega(stfem_problem p, int level, double *omega);

/* SystemMatrix::vmult / Tvmult, operators.h:94-152 (level -1: FP64 fine) */
int stfem_st_vmult(stfem_problem p, int level, void *dst, const void *src);
int stfem_st_tvmult(stfem_problem p, int level, void *dst, const void *src);
/* PreconditionVanka::vmult, gmg.h:476-514 */
int stfem_vanka_vmult(stfem_problem p, int level, void *dst, const void *src);
/* STMGTransferBlockMatrixFree::prolongate_and_add / restrict_and_add,
 * gmg.h:211-225: between `level`-1 (coarse) and `level` (fine) */
int stfem_prolongate_and_add(stfem_problem p, int level, void *dst, const void *src);
int stfem_restrict_and_add(stfem_problem p, int level, void *dst, const void *src);
/* GMG::vmult, gmg.h:846-859 (FP64 in/out; levels in their precision) */
int stfem_gmg_vmult(stfem_problem p, double *dst, const double *src);
/* SolverFGMRES::solve with ReductionControl(maxiter, abstol, reltol),
 * time_integrators.h:44, :174: x holds the initial guess.  its: iteration
 * count (last_step()); hist: residual history (hist_cap entries max). */
int stfem_fgmres_solve(stfem_problem p, double *x, const double *b, int *its, double *hist,
                       int hist_cap);
/* TimeIntegratorHeat::solve, time_integrators.h:157-180: rhs from prev
 * (rhs_matrix, operators.h:154-187) + source, extrapolated start, FGMRES,
 * Dirichlet nodes zeroed.  x: nt values per node; prev: one per node. */
int stfem_solve_slab(stfem_problem p, double *x, const double *prev, double time, int *its);
/* TimeIntegratorWave::solve, time_integrators.h:237-280 (problem=wave, the
 * first-order-in-time wave system of tp_01.cc:167-197): rhs from prev_u
 * (rhs_matrix) and prev_v (rhs_matrix_v) + source, extrapolated start,
 * FGMRES, Dirichlet nodes zeroed; then the velocities v of every block from
 * x, prev_u and prev_v (AixB, AixG, AixZ of time_integrators.h:219-233).
 * x, v: nt values per node; prev_u, prev_v: one per node. */
int stfem_solve_slab_wave(stfem_problem p, double *x, double *v, const double *prev_u,
                          const double *prev_v, double time, int *its);
/* ErrorCalculator::evaluate_error (exact_solution.h:218-299) for one slab
 * against the manufactured heat solution (source=0): Gauss(r+1) points in
 * space and time; out[0..2] = {squared L2 error, L-infinity error, squared
 * H1-seminorm error} of the slab starting at `time` (device x and prev as for
 * stfem_solve_slab). */
int stfem_slab_error(stfem_problem p, const double *x, const double *prev, double time,
                     double *out);
/* The heat driver of tp_01.cc:630-683 (TimeIntegratorHeat + ErrorCalculator)
 * from zero initial values to end_time: stats[0..6] = {slabs, total FGMRES
 * iterations, L2, Linf, H1-seminorm error of the whole run, fine nodes, time
 * dofs per slab}; its_per_slab (optional, cap entries) gets the iterations
 * of every slab.  Returns the first non-OK solver status. */
int stfem_run_heat(stfem_problem p, double *stats, int *its_per_slab, int cap);
/* support points of the fine nodes (host, n_nodes x dim, x fastest node
 * numbering): the Gauss-Lobatto points of FE_Q(p) mapped by each cell's Q1
 * map, as DoFTools::map_dofs_to_support_points gives them (tp_01.cc:415-438
 * evaluates solutions there) */
int stfem_node_coordinates(stfem_problem p, double *xyz);
/* host array h (one value per fine node): source f(t,x) = h(x) (int phi_i)
 * with unit time profile (source=2) */
int stfem_set_load_vector(stfem_problem p, const double *h);

/* ---- Adjoint of the heat slab sequence (dG time discretisation).  No
 * counterpart in the reference, which has only the forward path; built on
 * its slab system (TimeIntegratorHeat::solve, time_integrators.h:157-180)
 * and SystemMatrix::Tvmult (operators.h:124-152).  The sequence starts from
 * zero initial values; slab n covers [n*nts*tau, (n+1)*nts*tau).  Slab
 * vectors are node-major (nt values per node), one slab after the other.
 * Misfit J = 1/2 sum_n (x_n - d_n)^T (Alpha (x) M) (x_n - d_n): for dG,
 * Alpha is the time mass matrix, so J is half the squared L2(space-time)
 * norm of the mismatch. */
/* forward solutions of n_slabs slabs into traj (device, n_slabs*n_nodes*nt);
 * its: FGMRES iterations per slab (host, n_slabs) or NULL */
int stfem_forward_trajectory(stfem_problem p, int n_slabs, double *traj, int *its);
/* J for the observations obs (device, n_slabs*n_nodes*nt) */
int stfem_misfit(stfem_problem p, const double *obs, int n_slabs, double *misfit, int *its);
/* dJ/d(log kappa_c) for every fine cell into grad (device, n_cells) and J
 * into *misfit: forward sweep keeping only the slab start states, then
 * backward: each slab is solved again, then its adjoint
 *   A^T lambda_n = (Alpha (x) M) e_n + E^T R^T lambda_{n+1}
 * with FGMRES and the multigrid of the transposed system, and
 *   grad_c -= kappa_c sum_j lambda_{n,j}^T K_c (Alpha x_n)_j.
 * its: per slab the forward and the adjoint iterations (host, 2*n_slabs) or
 * NULL. */
int stfem_adjoint_gradient(stfem_problem p, const double *obs, int n_slabs, double *grad,
                           double *misfit, int *its);
/* new per-cell coefficient of the fine system (host, n_cells); the
 * multigrid preconditioners keep the coefficient they were built with */
int stfem_set_cell_coefficient(stfem_problem p, const double *kappa);
/* algorithmic HBM bytes of one application: what 0 fine vmult, 1 Vanka
 * sweep on `level`, 2 prolongate_and_add into `level`, 3 restrict_and_add
 * from `level` (source read, destination read and written) */
int stfem_bytes_model(stfem_problem p, int what, int level, double *bytes);
/* [t][node] <-> [node][t] on the device (double) */
int stfem_blocks_to_nodes(long n_nodes, int nt, const double *blocks, double *nodes);
int stfem_nodes_to_blocks(long n_nodes, int nt, const double *nodes, double *blocks);

/* ---- measurement (HIP events on the library stream) ----------------- */
/* average milliseconds of one application over `reps` after `warmup`:
 * what 0 fine FP64 vmult, 1 level vmult, 2 level Vanka, 3 GMG V-cycle,
 * 4 prolongate_and_add into `level`, 5 restrict_and_add from `level`, 6 the
 * sweep kernel of the level's factored Vanka alone (no node gather) */
int stfem_time_kernel(stfem_problem p, int what, int level, int warmup, int reps, double *ms);
/* TimerOutput scopes of the reference (printTiming): device time (ms) and
 * call counts of the phases {vmult, vanka, gmg, step} since creation or the
 * last reset: vmult = every SystemMatrix application (operators.h:97),
 * vanka = every smoother application (gmg.h:479), gmg = GMG::vmult
 * (gmg.h:850), step = one slab solve (tp_01.cc:632).  Needs the spec key
 * print_timing=1 (the V-cycle then launches eagerly; the totals are also
 * printed to stderr when the problem is destroyed).  reset != 0 zeroes the
 * totals after reading. */
int stfem_phase_timings(stfem_problem p, int reset, double *ms, long *calls);
/* runs `n_slabs` heat slabs of the configured run (from t = 0 and a zero
 * initial value to end_time in end_time / (nts * tau) slabs; after the last
 * slab of a run the next one starts again from t = 0 and zero); returns the
 * summed slab-solve time (ms, HIP events) and the summed FGMRES iterations.
 * Slabs that reach the iteration limit are timed and counted, and the run
 * goes on (status STFEM_ERR_NO_CONVERGENCE at the end); any other failure
 * stops it. */
int stfem_time_slabs(stfem_problem p, int n_slabs, double *ms, int *its);

/* ==== Object-level boundary =========================================
 * The reference's objects as separate handles, composed by the caller the
 * way tests/tp_01.cc:200-359 composes them: spaces (one DoFHandler +
 * coefficient per level), space-time operators (SystemMatrix), smoothers
 * (PreconditionVanka), two-level transfers, the GMG preconditioner and the
 * FGMRES solver.  Handles are owned by the caller (each has a _destroy) and
 * keep what they use alive (a GMG holds its levels and transfers; every
 * handle holds its context), so they may be destroyed in any order.  Device
 * vectors are node-major (x[node * nt + t]) in the object's precision
 * (precision 4: float, 8: double); calls are synchronous like the problem
 * API above; errors are reported through stfem_last_error(). */
typedef struct stfem_context_s  *stfem_context;
typedef struct stfem_space_s    *stfem_space;
typedef struct stfem_stop_s     *stfem_stop;
typedef struct stfem_vanka_s    *stfem_vanka;
typedef struct stfem_transfer_s *stfem_transfer;
typedef struct stfem_gmg_s      *stfem_gmg;
typedef struct stfem_fgmres_s   *stfem_fgmres;

/* device and stream (one MPI rank of the reference, tp_01.cc:740-750) */
int stfem_context_create(int device, stfem_context *out);
int stfem_context_destroy(stfem_context c);

/* The level spaces of tp_01.cc:83-90 and :262-275 (subdivided box refined
 * globally, FE_Q(p) on Gauss-Lobatto nodes, zero Dirichlet faces): one
 * space per refinement level 0..refine (coarsest first), built from the
 * mesh keys of a spec (dim, subdiv, lo, hi, refine, p, dirichlet,
 * coarse_coef) with the finest-level vertices and per-cell coefficient as
 * in stfem_problem_create; coarser levels evaluate the coefficient at their
 * quadrature points (MatrixFreeOperator::evaluate_coefficient,
 * operators.h:435-462).  cap: length of `spaces`; *n: spaces written. */
int stfem_space_hierarchy_create(stfem_context c, const char *spec, const double *vertices,
                                 const double *cell_coef, int cap, stfem_space *spaces, int *n);
int stfem_space_destroy(stfem_space s);
/* info[0..7] = dim, p, cells x, y, z, n_nodes, n_cells, Dirichlet face mask */
int stfem_space_info(stfem_space s, long *info);
/* host mask (n_nodes bytes): 1 for nodes constrained to zero (the
 * AffineConstraints of DoFTools::make_zero_boundary_constraints, tp_01.cc:92) */
int stfem_space_constrained(stfem_space s, unsigned char *mask);

/* SystemMatrix(timer, K, M, Alpha, Beta), operators.h:78-93: the space-time
 * operator Alpha (x) K + Beta (x) M with K the space's stiffness (with its
 * coefficient) and M its mass matrix, applied matrix-free
 * (MatrixFreeOperator, operators.h:355-562).  Alpha, Beta: row-major
 * nt_rows x nt_cols; nt_cols = 1 gives the rhs operator of
 * operators.h:154-187. */
int stfem_stop_create(stfem_space s, int nt_rows, int nt_cols, const double *alpha,
                      const double *beta, int precision, stfem_stop *out);
int stfem_stop_destroy(stfem_stop op);
/* info[0..3] = n_nodes, nt_rows, nt_cols, precision */
int stfem_stop_info(stfem_stop op, long *info);
/* SystemMatrix::vmult / Tvmult, operators.h:94-152 (square time matrices
 * for Tvmult) */
int stfem_stop_vmult(stfem_stop op, void *dst, const void *src);
int stfem_stop_tvmult(stfem_stop op, void *dst, const void *src);
/* No counterpart in the reference.  stfem_stop_vmult for callers that cannot
 * allocate device memory during the call (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation): the same result
 * with the operator's cell-major intermediate results in a fixed 64 MiB
 * device array of the library instead of the operator's own buffer.  Calls no
 * allocator.  Operators whose intermediate results exceed 64 MiB
 * (n_cells * (p+1)^dim * nt_rows * precision bytes) return STFEM_ERR_ARG.
 * The array is one per GPU and shared by every call on that GPU: calls that
 * overlap on one GPU (several threads or processes' streams) race. */
int stfem_stop_vmult_noalloc(stfem_stop op, void *dst, const void *src);
/* SystemMatrix::vmult_add(BlockVector&, const Vector&), operators.h:154-187:
 * dst (nt_rows values per node) += (Alpha(:,0) (x) K + Beta(:,0) (x) M) src
 * (one value per node); needs nt_cols = 1 */
int stfem_stop_vmult_add(stfem_stop op, void *dst, const void *src);

/* PreconditionVanka(timer, K, M, SP, Alpha, Beta, dof_handler),
 * gmg.h:409-474: cell patches of the space, patch blocks of the assembled K
 * and M; form 0: explicit inverses (gauss_jordan, gmg.h:470), 1: the same
 * inverse factored through the generalized eigenproblem of (K_c, M_c);
 * device_setup as the spec key vanka_device_setup (3 default). */
int stfem_vanka_create(stfem_space s, int nt, const double *alpha, const double *beta,
                       int precision, int form, int device_setup, stfem_vanka *out);
int stfem_vanka_destroy(stfem_vanka v);
/* PreconditionVanka::vmult, gmg.h:476-514 */
int stfem_vanka_apply(stfem_vanka v, void *dst, const void *src);

/* MGTwoLevelTransferST(type, r, n_timesteps_at_once, restrict_is_transpose,
 * mg_type), gmg.h:86-115: kind 1 tau (between nts and nts/2 steps of
 * degree r), 2 k (between degree r and r - 1); on the nodes of space s. */
int stfem_time_transfer_create(stfem_space s, int type, int r, int nts, int kind,
                               int restrict_is_transpose, int precision, stfem_transfer *out);
/* MGTwoLevelBlockTransfer (gmg.h:35-75) over deal.II's MGTwoLevelTransfer
 * between two spaces one refinement apart (gmg.h:371-380), nt blocks */
int stfem_space_transfer_create(stfem_space fine, stfem_space coarse, int nt, int precision,
                                stfem_transfer *out);
int stfem_transfer_destroy(stfem_transfer t);
/* prolongate_and_add / restrict_and_add, gmg.h:117-135 and :56-74 */
int stfem_transfer_prolongate_and_add(stfem_transfer t, void *dst, const void *src);
int stfem_transfer_restrict_and_add(stfem_transfer t, void *dst, const void *src);

/* GMG(..., mg_operators, mg_smoother (Vanka), transfers, ...),
 * gmg.h:666-790: levels 0 (coarsest) .. n_levels-1, transfers[l] between
 * l-1 and l (transfers[0] unused), all in one precision.  options: spec keys
 * smoothing_steps, variable, estimate_relaxation, eig_iters, eig_safety,
 * exact_single_cell, use_graph, dense_coarse_dofs, fuse_residual; omega:
 * one relaxation parameter per level or NULL (estimated as
 * GMG::estimate_relaxation does, gmg.h:792-844). */
int stfem_gmg_create(int n_levels, const stfem_stop *ops, const stfem_vanka *smoothers,
                     const stfem_transfer *transfers, const char *options, const double *omega,
                     stfem_gmg *out);
int stfem_gmg_destroy(stfem_gmg g);
int stfem_gmg_omega(stfem_gmg g, int level, double *omega);
/* GMG::vmult, gmg.h:846-859: FP64 in and out, the levels in their precision */
int stfem_gmg_apply(stfem_gmg g, double *dst, const double *src);

/* SolverFGMRES with ReductionControl(maxiter, abstol, reltol) and a basis
 * of `basis` vectors (time_integrators.h:44, :113) for systems of n FP64
 * unknowns */
int stfem_fgmres_create(stfem_context c, long n, int maxiter, double abstol, double reltol,
                        int basis, int orthogonalization, stfem_fgmres *out);
int stfem_fgmres_destroy(stfem_fgmres s);
/* SolverFGMRES::solve(A, x, b, P), time_integrators.h:174: A an FP64
 * operator, P a GMG or NULL (no preconditioner); x holds the start vector.
 * Returns 0, STFEM_ERR_NO_CONVERGENCE or STFEM_ERR_NOT_FINITE; its and hist
 * (residual history, hist_cap entries max) as stfem_fgmres_solve. */
int stfem_fgmres_run(stfem_fgmres s, stfem_stop A, stfem_gmg P, double *x, const double *b,
                     int *its, double *hist, int hist_cap);

#ifdef __cplusplus
}
#endif
#endif
