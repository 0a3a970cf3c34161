// Host C++ surface of the MI355X space-time multigrid slab solver.
//
// Class names and call semantics follow nlsmrg/dealii-stfem:
//   SystemMatrix ............ tests/include/operators.h:71-256
//   PreconditionVanka ....... tests/include/gmg.h:402-540
//   MGTwoLevelTransferST .... tests/include/gmg.h:77-136
//   MGTwoLevelBlockTransfer . tests/include/gmg.h:35-75 (space h-transfer)
//   GMG ..................... tests/include/gmg.h:652-906
//   SolverFGMRES + ReductionControl (deal.II) as used in
//                             tests/include/time_integrators.h:21-120
//   TimeIntegratorHeat ...... tests/include/time_integrators.h:123-181
// Vectors live in device memory in node-major layout: the nt time values of
// a node are contiguous (x[node * nt + t]).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "mesh.h"
#include "spec.h"
#include "time_fe.h"

namespace stfem
{
  // ------------------------------------------------------------- device mem
  template <typename T>
  class DBuf
  {
  public:
    DBuf() = default;
    explicit DBuf(size_t n) { alloc(n); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept
      : p_(o.p_)
      , n_(o.n_)
    {
      o.p_ = nullptr;
      o.n_ = 0;
    }
    DBuf &operator=(DBuf &&o) noexcept
    {
      if (this != &o)
        {
          release();
          p_   = o.p_;
          n_   = o.n_;
          o.p_ = nullptr;
          o.n_ = 0;
        }
      return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t n)
    {
      release();
      n_ = n;
      if (n_)
        STFEM_HIP(hipMalloc(&p_, n_ * sizeof(T)));
    }
    void release()
    {
      if (p_)
        (void)hipFree(p_);
      p_ = nullptr;
      n_ = 0;
    }
    void upload(const T *h, size_t n, hipStream_t s)
    {
      if (n_ < n)
        alloc(n);
      if (n)
        STFEM_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void zero(hipStream_t s)
    {
      if (n_)
        STFEM_HIP(hipMemsetAsync(p_, 0, n_ * sizeof(T), s));
    }
    T       *get() { return p_; }
    const T *get() const { return p_; }
    size_t   size() const { return n_; }

  private:
    T     *p_ = nullptr;
    size_t n_ = 0;
  };

  // ---------------------------------------------------------- phase timers
  // Device time per phase, the TimerOutput scopes of the reference: "vmult"
  // (SystemMatrix::vmult, operators.h:97; every operator application of a
  // level or of the fine system), "vanka" (PreconditionVanka::vmult,
  // gmg.h:479; here every relaxation sweep, which applies the smoother),
  // "gmg" (GMG::vmult, gmg.h:850) and "step" (one slab solve, tp_01.cc:632).
  // Scopes may nest.  Measured with HIP events on the library stream and
  // summed when read.
  class PhaseTimers
  {
  public:
    enum Phase
    {
      vmult    = 0,
      vanka    = 1,
      gmg      = 2,
      step     = 3,
      n_phases = 4
    };
    ~PhaseTimers();
    hipEvent_t start(hipStream_t s);
    void       stop(int phase, hipEvent_t a, hipStream_t s);
    // totals (ms) and call counts so far
    void       read(double *ms, long *calls);
    void       reset();
    static const char *name(int phase);

  private:
    struct Pending
    {
      int        phase;
      hipEvent_t a, b;
    };
    hipEvent_t           get();
    void                 resolve();
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    double               ms_[n_phases]    = {};
    long                 calls_[n_phases] = {};
  };

  // --------------------------------------------------------------- context
  struct Context
  {
    int              device = 0;
    hipStream_t      stream = nullptr;
    DBuf<double>     partial;  // reduction scratch
    DBuf<double>     scalars;  // device scalars (Hessenberg column, norms)
    std::unique_ptr<PhaseTimers> timers; // set: phases are timed (print_timing)
    explicit Context(int dev);
    ~Context();
    void sync();
  };

  // times the enclosing block as one call of `phase` when the context's
  // timers are on
  class PhaseScope
  {
  public:
    PhaseScope(const Context &ctx, int phase)
      : t_(ctx.timers.get())
      , s_(ctx.stream)
      , phase_(phase)
    {
      if (t_)
        a_ = t_->start(s_);
    }
    ~PhaseScope()
    {
      if (t_)
        t_->stop(phase_, a_, s_);
    }
    PhaseScope(const PhaseScope &)            = delete;
    PhaseScope &operator=(const PhaseScope &) = delete;

  private:
    PhaseTimers *t_;
    hipStream_t  s_;
    int          phase_;
    hipEvent_t   a_ = nullptr;
  };

  // ------------------------------------------------------ SystemMatrix
  template <typename T>
  class SystemMatrix
  {
  public:
    // two_pass: cell results through a cell-major scratch buffer and a
    // per-node gather (default); otherwise one pass with atomic adds
    SystemMatrix(Context &ctx, const MeshLevel &m, const DMat &Alpha, const DMat &Beta,
                 bool two_pass = true);
    // dst = (Alpha x K + Beta x M) src   (operators.h:94-122)
    void vmult(T *dst, const T *src) const;
    // the same with the cell-major results in `cells` (at least
    // cell_buffer_bytes()) instead of the operator's own buffer
    void vmult_with_cells(T *dst, const T *src, T *cells) const;
    size_t cell_buffer_bytes() const { return st_cells_size(grid_, nt_out_) * sizeof(T); }
    // the cell pass of vmult alone, into cell_buffer() (two-pass form)
    void vmult_cells(const T *src) const;
    const T *cell_buffer() const { return cells_.get(); }
    // dst = (Alpha^T x K + Beta^T x M) src   (operators.h:124-152)
    void Tvmult(T *dst, const T *src) const;
    // dst(all blocks) += (Alpha(:,0) x K + Beta(:,0) x M) src(one block)
    // for an nt x 1 weight pair (operators.h:154-187)
    void vmult_add_nx1(T *dst, const T *src) const;
    // dst = rhs - (Alpha x K + Beta x M) src (the residual of the V-cycle)
    void residual(T *dst, const T *rhs, const T *src) const;
    // replaces the coefficient by one value per cell (host array, n_cells)
    void set_cell_coefficient(const double *kappa);
    // output time components per cell pass of the kernel (1..4): fewer
    // means fewer registers and more passes over the inputs
    void set_max_jb(int j) { max_jb_ = j; }
    // grids of fewer cell workgroups than this compute one output
    // component per workgroup (StOpArgs::split_below)
    void set_split_below(int b) { split_below_ = b; }
    // 3D cell pass: 1 wavefront-local cells, 0 workgroup cells (StOpArgs::kernel)
    void set_op_kernel(int k) { op_kernel_ = k; }
    long        n_nodes() const { return grid_.n_nodes; }
    int         n_blocks() const { return nt_out_; }
    const double *vertices() const { return verts_.get(); }
    long        m() const { return grid_.n_nodes * nt_out_; }
    const GridDesc &grid() const { return grid_; }
    const DMat &alpha() const { return A_; }
    const DMat &beta() const { return B_; }
    // algorithmic bytes of one vmult (vectors, coefficient, vertices)
    double      vmult_bytes() const;

  private:
    void apply(T *dst, const T *src, const DBuf<T> &a, const DBuf<T> &b, int mode,
               const T *rhs = nullptr, T *cells = nullptr) const;
    Context     &ctx_;
    GridDesc     grid_;
    DBuf<double> verts_;
    int          coef_mode_;
    T            coef_value_;
    DBuf<T>      coef_;
    DMat         A_, B_;
    int          nt_in_, nt_out_;
    int          max_jb_ = 4, split_below_ = 0, op_kernel_ = 1;
    DBuf<T>      dA_, dB_, dAt_, dBt_;
    mutable DBuf<T> cells_; // cell-major results of the two-pass form
  };

  // ------------------------------------------------------ PreconditionVanka
  template <typename T>
  class PreconditionVanka
  {
  public:
    enum class Form
    {
      explicit_inverse,
      factored
    };
    // device_setup: the factored form is set up on the GPU (element and
    // patch matrices, batched generalized eigensolver: 1 Jacobi, 2 divide
    // and conquer, 3 by patch size); 0, and the explicit inverse: on the host
    PreconditionVanka(Context &ctx, const MeshLevel &m, const Shape1D &sh, const DMat &Alpha,
                      const DMat &Beta, Form form, int device_setup = 3);
    // dst = sum_c (1/valence) B_c^{-1} src|_c   (gmg.h:476-514)
    void   vmult(T *dst, const T *src) const;
    // u += omega * Vanka(rhs - ax); ax may be null; assign: u = omega *
    // Vanka(rhs - ax) (u need not be initialised).  One relaxation sweep of
    // PreconditionRelaxation with this preconditioner.
    void   relax(T *u, const T *rhs, const T *ax, T omega, bool assign = false) const;
    // the same update with A x summed from the cell buffer of the level
    // operator's cell pass (SystemMatrix::vmult_cells), when
    // fused_residual_ok(): one node gather fewer per sweep
    bool   fused_residual_ok() const;
    void   relax_opc(T *u, const T *rhs, const T *opc, T omega) const;
    // the sweep kernel of vmult alone (factored form with the cell-major
    // buffer): corrections of src into the cell buffer, no node gather
    void   cell_pass(const T *src) const;
    long   n_cells() const { return grid_.n_cells; }
    long   n_nodes() const { return grid_.n_nodes; }
    int    n_blocks() const { return nt_; }
    Form   form() const { return form_; }
    // factored sweep kernel (0 automatic, 2 LDS-staged, 3 register-tiled)
    void   set_variant(int v) { variant_ = v; }
    double sweep_bytes() const;
    size_t storage_bytes() const { return (binv_.size() + S_.size() + lam_.size()) * sizeof(T); }

  private:
    void     setup_factored_on_device(const MeshLevel &m, const Shape1D &sh, const DMat &Alpha,
                                      const DMat &Beta, int method);
    Context &ctx_;
    GridDesc grid_;
    int      nt_;
    Form     form_;
    DBuf<T>  binv_, S_, lam_; // explicit inverses; eigenvectors and eigenvalues
    DBuf<T>  ta_, tb_;        // the level's time matrices (factored form)
    mutable DBuf<T> cells_; // cell-major corrections of the two-pass sweep
    int             variant_ = 0;
  };

  // ------------------------------------------------------ transfers
  template <typename T>
  class MGTwoLevelTransferST
  {
  public:
    // gmg.h:86-115 (restrict_is_transpose: R = P^T, else the L2
    // restriction / projection matrices); `overwrite`
    // reproduces the assigning transfer of tests/transfer_01.cc:42-83.
    MGTwoLevelTransferST(Context &ctx, TimeStepType type, int r, int nts, TimeMGType kind,
                         long n_nodes, bool overwrite = false, bool restrict_is_transpose = true);
    void prolongate_and_add(T *dst, const T *src) const;
    void restrict_and_add(T *dst, const T *src) const;
    int  nt_fine() const { return P_.m; }
    int  nt_coarse() const { return P_.n; }
    bool overwrites() const { return overwrite_; }
    const DMat &prolongation() const { return P_; }

  private:
    Context &ctx_;
    DMat     P_;
    long     n_nodes_;
    bool     overwrite_;
    DBuf<T>  dP_, dR_;
  };

  template <typename T>
  class MGTwoLevelBlockTransfer
  {
  public:
    MGTwoLevelBlockTransfer(Context &ctx, const GridDesc &fine, const GridDesc &coarse,
                            const Shape1D &sh, int nt);
    void prolongate_and_add(T *dst, const T *src) const;
    void restrict_and_add(T *dst, const T *src) const;

  private:
    struct Ell
    {
      int          rows = 0, K = 0;
      DBuf<int>    cnt, col;
      DBuf<T>      w;
    };
    Context &ctx_;
    GridDesc fine_, coarse_;
    int      nt_;
    Ell      P_[3], R_[3];
    DBuf<T>  tmp1_, tmp2_;
  };

  // One two-level transfer: time (tau/k) or space (h), as the std::variant
  // of TwoLevelTransferOperator (gmg.h:138-177).
  template <typename T>
  struct TwoLevelTransfer
  {
    std::shared_ptr<MGTwoLevelTransferST<T>>    time;
    std::shared_ptr<MGTwoLevelBlockTransfer<T>> space;
    void prolongate_and_add(T *dst, const T *src) const
    {
      time ? time->prolongate_and_add(dst, src) : space->prolongate_and_add(dst, src);
    }
    void restrict_and_add(T *dst, const T *src) const
    {
      time ? time->restrict_and_add(dst, src) : space->restrict_and_add(dst, src);
    }
    // the prolongation replaces dst instead of adding to it
    bool prolongation_overwrites() const { return time && time->overwrites(); }
  };

  // ------------------------------------------------------------ GMG
  // Level data of the hierarchy (precision T).
  template <typename T>
  struct GMGLevel
  {
    std::shared_ptr<SystemMatrix<T>>      A;
    std::shared_ptr<PreconditionVanka<T>> V;
    double                                omega = 1.0;
    int                                   r = 0, nts = 1;
    TimeMGType                            type_to_here = TimeMGType::none;
  };

  // Geometric space-time multigrid V-cycle used as FGMRES preconditioner
  // (gmg.h:652-906): Vanka relaxation smoother with variable step counts,
  // coarse level solved by the smoother, FP64 <-> T conversion at the
  // boundary (gmg.h:846-859).
  template <typename T>
  class GMG
  {
  public:
    // omega: relaxation parameter per level (nullptr: estimated as
    // GMG::estimate_relaxation does, or 1 without estimate_relaxation)
    GMG(Context &ctx, std::vector<GMGLevel<T>> &&levels,
        std::vector<TwoLevelTransfer<T>> &&transfers, const GMGOptions &opt,
        const std::vector<double> *omega = nullptr);
    ~GMG();
    void   vmult(double *dst, const double *src) const;
    // one V-cycle on level vectors (def[L] given, sol[L] returned)
    void   vcycle() const;
    double estimate_omega(int level) const;
    int    n_levels() const { return int(lv_.size()); }
    const GMGLevel<T> &level(int l) const { return lv_[l]; }
    TwoLevelTransfer<T> &transfer(int l) { return tr_[l]; }
    T *level_defect(int l) const { return def_[l].get(); }
    T *level_solution(int l) const { return sol_[l].get(); }
    // level whose V-cycle is one dense matrix-vector product (-1: none)
    int dense_level() const { return dense_level_; }

  private:
    void setup_dense_coarse();
    void relax(int l, T *u, const T *rhs, bool first) const;
    void relax_omega(int l, T *u, const T *rhs, bool first, T omega) const;
    bool exact_level(int l) const;
    T    exact_omega(int l, int sweeps) const;
    void smoother_apply(int l, T *u, const T *rhs) const;
    void smoother_smooth(int l, T *u, const T *rhs) const;
    void level_v_step(int l) const;
    int  steps(int l) const;

    Context                         &ctx_;
    std::vector<GMGLevel<T>>         lv_;
    std::vector<TwoLevelTransfer<T>> tr_;
    GMGOptions                       opt_;
    mutable std::vector<DBuf<T>>     sol_, def_, t1_, t2_;
    mutable hipGraph_t               graph_ = nullptr;
    mutable hipGraphExec_t           exec_  = nullptr;
    bool                             use_graph_   = false;
    int                              dense_level_ = -1;
    long                             dense_ld_    = 0;
    DBuf<T>                          dense_rows_; // row-major, rows padded to dense_ld_
  };

  // ------------------------------------------------------------ FGMRES
  // The identity as FGMRES preconditioner (the solver without multigrid).
  struct IdentityPreconditioner
  {
    Context *ctx;
    long     n;
    void     vmult(double *dst, const double *src) const
    {
      STFEM_HIP(hipMemcpyAsync(dst, src, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice,
                               ctx->stream));
    }
  };

  // Right-preconditioned flexible GMRES (deal.II SolverFGMRES semantics:
  // modified Gram-Schmidt, projected least-squares solve once two basis
  // vectors exist, ReductionControl on the projected residual).
  class SolverFGMRES
  {
  public:
    SolverFGMRES(Context &ctx, long n, const SolverControlData &c);
    template <typename Op, typename Prec>
    int solve(const Op &A, const Prec &P, double *x, const double *b);
    int                        last_step() const { return last_step_; }
    const std::vector<double> &history() const { return hist_; }

  private:
    Context                  &ctx_;
    long                      n_;
    SolverControlData         c_;
    std::vector<DBuf<double>> v_, z_;
    DBuf<double>              aux_;
    DBuf<double>              gs_part_, gs_dots_, gs_h_, gs_G_; // one-reduction Gram-Schmidt
    int                       last_step_ = 0;
    std::vector<double>       hist_;
  };

  // ------------------------------------------------------------ problem
  class ProblemBase
  {
  public:
    virtual ~ProblemBase() = default;
    ProblemSpec spec;
    double      tau = 0;
    int         nt  = 1;
    virtual int  n_levels() const = 0;
    virtual void level_info(int l, long *info) const = 0;
    virtual double level_omega(int l) const = 0;
    virtual void level_vmult(int l, void *dst, const void *src, bool transpose) = 0;
    virtual void level_vanka(int l, void *dst, const void *src) = 0;
    virtual void level_vanka_cells(int l, const void *src) = 0;
    virtual void level_prolongate(int l, void *dst, const void *src) = 0;
    virtual void level_restrict(int l, void *dst, const void *src) = 0;
    virtual void fine_vmult(double *dst, const double *src, bool transpose) = 0;
    virtual void gmg_vmult(double *dst, const double *src) = 0;
    virtual int  fgmres(double *x, const double *b, int *its, std::vector<double> *hist) = 0;
    virtual int  solve_slab(double *x, const double *prev, double time, int *its) = 0;
    // TimeIntegratorWave::solve: displacement x and velocity v of one slab
    // from the previous slab's last displacement and velocity
    virtual int  solve_slab_wave(double *x, double *v, const double *prev_u,
                                 const double *prev_v, double time, int *its) = 0;
    // ErrorCalculator::evaluate_error of one slab: out = {L2^2, Linf, H1semi^2}
    virtual void slab_error(const double *x, const double *prev, double time, double *out) = 0;
    // the heat time loop of tp_01.cc:630-683 from zero initial values:
    // stats = {slabs, total iterations, L2, Linf, H1semi, n_nodes, nt}
    virtual int  run_heat(double *stats, int *its_per_slab, int cap) = 0;
    virtual void set_load_vector(const double *h_host) = 0;
    // fine-level node coordinates (n_nodes x dim, host)
    virtual void node_coordinates_host(std::vector<double> &xyz) const = 0;
    // ---- adjoint of the heat slab sequence (dG time), misfit
    //   J = 1/2 sum_n (x_n - d_n)^T (Alpha (x) M) (x_n - d_n)
    // over n_slabs slabs from zero initial values; obs and traj hold the
    // node-major slab vectors one after the other (device).
    virtual void forward_trajectory(int n_slabs, double *traj, int *its) = 0;
    virtual double misfit(const double *obs, int n_slabs, int *its) = 0;
    // dJ/d(log kappa_c) into grad (device, n_cells), misfit into *J; its:
    // per slab the forward and the adjoint FGMRES iterations (2*n_slabs)
    virtual void adjoint_gradient(const double *obs, int n_slabs, double *grad, double *J,
                                  int *its) = 0;
    // new per-cell coefficient of the fine system (host, n_cells); the
    // multigrid preconditioners keep the coefficient they were built with
    virtual void set_cell_coefficient(const double *kappa) = 0;
    virtual long n_nodes() const = 0;
    virtual Context &context() = 0;
    virtual double fine_vmult_bytes() const = 0;
    virtual double vanka_sweep_bytes(int l) const = 0;
  };

  // dst = op src with the operator's cell-major results in a fixed device
  // array of the library (no allocation); throws if they do not fit
  template <typename T>
  void   vmult_noalloc(const SystemMatrix<T> &op, int device, T *dst, const T *src);
  size_t noalloc_scratch_bytes();

  std::unique_ptr<ProblemBase> make_problem(const std::string &spec, const double *vertices,
                                            const double *kappa, int device);
} // namespace stfem
