// Host-side launchers of the gfx950 kernels.  All pointers are device
// pointers; all launches are asynchronous on the given stream.
#pragma once

#include "common.h"

#include <string>
#include <type_traits>
#include <vector>

namespace stfem
{
  // ------------------------------------------------------------ shape data
  // 1D tables of the FE_Q(p) element on Gauss-Lobatto nodes at the
  // Gauss(p+1) points: S[q][a] = phi_a(x_q), D[q][q'] = derivative of the
  // Lagrange polynomial on the Gauss points (collocation), W[q] = weight.
  // Uploaded once per device into constant memory for p = 1..MAX_N-1.
  void upload_shape_tables(const double *S, const double *D, const double *W, const double *X,
                           int N);

  // ------------------------------------------------------- space-time op
  // dst[node][j] (+)= sum_i Alpha(j,i) (K src_i)[node] + Beta(j,i) (M src_i)[node]
  // K: stiffness int kappa grad phi . grad phi (Q1 mapping), M: mass.
  // Dirichlet nodes are read as zero; their dst entries are set to zero
  // (or left unchanged when accumulating).
  template <typename T>
  struct StOpArgs
  {
    GridDesc      g;
    const double *vertices;  // (nc+1)^dim vertex coordinates, dim per vertex
    int           coef_mode; // 0: constant coef_value, 1: per cell, 2: per (cell, q)
    T             coef_value;
    const T      *coef;
    const T      *alpha;     // nt_out x nt_in (stiffness weights), device
    const T      *beta;      // nt_out x nt_in (mass weights), device
    int           nt_in, nt_out;
    int           max_jb = 4; // output components per cell pass (1..4)
    // grids of fewer cell workgroups than this compute one output component
    // per workgroup (grid.y = nt_out): more workgroups in flight and a
    // shorter chain per workgroup on small levels, at the price of
    // evaluating the inputs once per component (0: off)
    int           split_below = 0;
    // 3D two-pass form: 1 (default) the cells of a workgroup are handled
    // by independent wavefronts (lines redistributed through LDS within a
    // wavefront), 0 by the whole workgroup (lines in registers, the other
    // directions swept through LDS between workgroup barriers)
    int           kernel = 1;
  };
  // cells != null: two-pass form with the cell-major scratch `cells` of
  // st_cells_size() values; cells == null: one pass with atomic adds.
  // mode 0: dst = A src;  1: dst += A src;  2: dst = rhs - A src (two-pass
  // form only); 3: only the cell pass into `cells` (dst unused), for a
  // consumer that sums the node values itself.
  template <typename T>
  void launch_st_apply(const StOpArgs<T> &a, const T *src, T *dst, T *cells, int mode,
                       const T *rhs, hipStream_t s);
  inline size_t st_cells_size(const GridDesc &g, int nt_out)
  {
    size_t nc = 1;
    for (int d = 0; d < g.dim; ++d)
      nc *= size_t(g.p + 1);
    return size_t(g.n_cells) * nc * size_t(nt_out);
  }

  // ------------------------------------------------------------ Vanka
  // u += omega * sum_cells (1/valence) * scatter( Binv_c * gather(rhs - ax) )
  // (ax may be null) with explicit per-cell inverses stored column-major,
  // nb = nt * (p+1)^dim.
  template <typename T>
  void launch_vanka_explicit(const GridDesc &g, int nt, const T *binv, const T *rhs, const T *ax,
                             T *u, T omega, hipStream_t s);
  // row stride of the factored Vanka's S blocks: rows padded to 16 bytes
  inline int vanka_s_ld(int nloc, size_t tsize)
  {
    const int vw = int(16 / tsize);
    return (nloc + vw - 1) / vw * vw;
  }
  // values per cell of the cell-block layout in which the register-tiled
  // factored sweep leaves its per-cell results: n_c * nt rounded up to whole
  // 128-byte lines (the size per cell of the sweep's cell buffer)
  inline constexpr long vanka_cell_block(int nloc, int nt, size_t tsize)
  {
    const long v = 128 / long(tsize);
    return (long(nloc) * nt + v - 1) / v * v;
  }
  // Factored form: Binv_c = (I x S_c) blockdiag_k (W_{c,k}) (I x S_c^T),
  // S_c (n_c x n_c, row-major with the row stride vanka_s_ld(n_c, sizeof(T)),
  // padding zero), W_{c,k} = (lam_{c,k} Alpha + Beta)^{-1} applied from the
  // eigenvalues lam (n_c per cell) and the level's time matrices ta = Alpha,
  // tb = Beta (nt x nt row-major, device).
  // cells (n_cells * vanka_cell_block(n_c, nt, sizeof(T)) values) non-null:
  // two-pass form for the
  // compile-time patch shapes, mode 0 u = correction, 1 u += correction;
  // otherwise (and for the other shapes) u += correction with atomics
  // (mode 0 then zero-fills u first).  variant: 0 automatic, 2 the
  // LDS-staged kernel, 3 the register-tiled kernel (patches of at most 128
  // nodes), 4 its software-pipelined persistent form (FP32); shapes a
  // variant does not cover fall back to the others.
  // mode 0: u = omega V r, 1: u += omega V r, 3 (two-pass form): only the
  // cell pass into `cells` (u unused; for timing the sweep kernel alone).
  // opc non-null (two-pass form only, shapes with
  // vanka_fused_residual_supported): A x at the patch nodes is summed from
  // the operator's cell-major buffer opc (StOpArgs's level, nt components)
  // instead of read from ax, which saves the operator's node gather.
  template <typename T>
  void launch_vanka_factored(const GridDesc &g, int nt, const T *S, const T *lam, const T *ta,
                             const T *tb, const T *rhs, const T *ax, T *u, T omega, T *cells,
                             int mode, int variant, hipStream_t s, const T *opc = nullptr);
  template <typename T>
  bool vanka_fused_residual_supported(const GridDesc &g, int nt);
  // Node gather of a cell-major buffer (cell_slot layout, nto components;
  // cs > 0: cell c's values contiguous from c * cs on, local node index
  // a + N b + N^2 c, components innermost): mode 0 dst = sum, 1 dst += sum,
  // 2 dst = rhs - sum; skip_dirichlet: Dirichlet nodes are left out (zero
  // for modes 0, 2, unchanged for 1).
  template <typename T>
  void launch_gather_cells(const GridDesc &g, int nto, const T *cells, T *dst, const T *rhs,
                           int mode, bool skip_dirichlet, hipStream_t s, long cs = 0);
  // p[0..n) = 0 with a kernel (safe to record into a HIP graph)
  template <typename T>
  void launch_fill_zero(T *p, size_t n, hipStream_t s);

  // ------------------------------------------------------------ transfers
  // Per node: dst[node][a] (= or +=) sum_b M(a,b) src[node][b].
  template <typename T>
  void launch_time_transfer(long n_nodes, const T *M, int nt_dst, int nt_src, const T *src,
                            T *dst, bool add, hipStream_t s);
  // 1D banded operator applied along direction d of a node array with
  // dims n[0..2] and nt interleaved components.  Row o of the operator has
  // cnt[o] entries (col, w) stored at o*K.
  template <typename T>
  void launch_sweep(const T *in, T *out, int nt, int d, const int n_in[3], int m_out,
                    const int *cnt, const int *col, const T *w, int K, bool add, hipStream_t s);

  // ------------------------------------------------------------ BLAS-1
  // Deterministic two-stage reductions.  `partial` needs reduce_slots()
  // doubles.
  int  reduce_slots();
  // out[0] (op)= dot(x, y)  (op: 0 assign, 1 assign sqrt)
  template <typename T>
  void launch_dot(long n, const T *x, const T *y, double *partial, double *out, int op,
                  hipStream_t s);
  // stage 1 only: partial[0..reduce_slots()) of dot(x, y)
  template <typename T>
  void launch_dot_partials(long n, const T *x, const T *y, double *partial, hipStream_t s);
  // One modified Gram-Schmidt step: c = sum of the partials `pin` (stored to
  // *hout when hout is not null), x += factor * c * y, then the partials of
  // dot(x, z) (z null: dot(x, x)) into `pout` (pout != pin).
  void launch_axpy_dot_chain(long n, double *x, const double *pin, double factor, const double *y,
                             const double *z, double *pout, double *hout, hipStream_t s);
  // stage 2 on the host, in the same order as the device stage 2
  double sum_partials(const double *partial_host);
  // Gram-Schmidt step j of a Krylov basis v[0..j] (j < 64) with one
  // reduction pass: the dots of x and v_j with the basis, the coefficients
  // h[0..j] of modified Gram-Schmidt from them through the Gram matrix G
  // (column j stored, leading dimension ldg > j), x -= sum h_i v_i, and the
  // partial sums of (x, x) into pnorm.  partial: (2j + 1) * reduce_slots()
  // doubles, dots: 2j + 1 doubles, h: j + 1 doubles (device).
  void launch_gs_step(long n, int j, const double *const *v, double *x, double *G, int ldg,
                      double *partial, double *dots, double *h, double *pnorm, hipStream_t s);
  template <typename T>
  void launch_axpby(long n, T a, const T *x, T b, T *y, hipStream_t s); // y = a x + b y
  void launch_multi_axpy(long n, int m, const double *const *z, const double *coef,
                         double *x, hipStream_t s);                      // x += sum coef_j z_j
  template <typename TD, typename TS>
  void launch_convert(long n, const TS *src, TD *dst, hipStream_t s);
  // [t][node] <-> [node][t]
  template <typename T>
  void launch_blocks_to_interleaved(long n_nodes, int nt, const T *blocks, T *inter,
                                    hipStream_t s);
  template <typename T>
  void launch_interleaved_to_blocks(long n_nodes, int nt, const T *inter, T *blocks,
                                    hipStream_t s);
  // zero the Dirichlet nodes of an interleaved vector
  template <typename T>
  void launch_zero_dirichlet(const GridDesc &g, int nt, T *x, hipStream_t s);

  // ------------------------------------------------------- Vanka setup
  // Element mass and stiffness matrices (row-major nl x nl, FP64) of cells
  // [c0, c0 + nc): Gauss(p+1) rule (qx, qw on [0,1]) with phi_a(x_q) = S and
  // phi_a'(x_q) = Sd ([q][a]), Q1 mapping from the vertices X, coefficient
  // as in StOpArgs (FP64).
  void launch_element_matrices(const GridDesc &g, const double *X, int coef_mode,
                               double coef_value, const double *coef, const double *S,
                               const double *Sd, const double *qx, const double *qw, long c0,
                               long nc, double *Me, double *Ke, hipStream_t s);
  // Patch matrices K[I_c,I_c], M[I_c,I_c] of the assembled matrices for
  // cells [p0, p0 + np) (column-major nl x nl; constrained rows/columns
  // reduced to the diagonal) from the element matrices of cells
  // [e0, e0 + ne), which must contain every neighbour of the patches.
  void launch_patch_matrices(const GridDesc &g, long p0, long np, long e0, const double *Me,
                             const double *Ke, double *Mp, double *Kp, hipStream_t s);
  // Batched generalized eigensolver of np patches of order n <= 128 (FP64):
  // K_c z = lambda M_c z, z^T M_c z = 1.  Mp, Kp: column-major n x n per
  // patch, symmetric (Mp is overwritten with the Cholesky factor, Kp with
  // the eigenvectors, eigenvector k in column k); lam: the eigenvalues (n per
  // patch, unsorted); ee, hh: n per patch of work space; rot: the rotation
  // log, patch_eigen_rotation_cap(n) entries per patch; chains:
  // patch_eigen_chain_cap(n) per patch; counts: two per patch.  info[c] = 0
  // on success, 1 for a mass matrix that is not positive definite, 2 for an
  // eigenvalue iteration that did not converge, 3 for a log that would
  // overflow.
  long   patch_eigen_rotation_cap(int n);
  int    patch_eigen_chain_cap(int n);
  size_t patch_eigen_lds(int n);
  void   launch_patch_eigen(int n, long np, double *Mp, double *Kp, double *lam, double *ee,
                            double *hh, double2 *rot, int2 *chains, int *counts, int *info,
                            hipStream_t s);
  // S_c (row-major with the row stride vanka_s_ld, eigenvector k = column k)
  // from the column-major eigenvectors Z and the eigenvalues in the level
  // precision (lam_out, n_c per cell); sets *singular when a time block
  // lam_k Alpha + Beta cannot be inverted (nt <= 8).
  template <typename T>
  void launch_pack_factors(long ncell, int nl, int nt, const double *Z, const double *lam,
                           const double *Alpha, const double *Beta, T *S, T *lam_out,
                           int *singular, hipStream_t s);

  // ------------------------------------------------------- adjoint
  // grad[c] -= kappa[c] * sum_j lam_j^T K_c w_j with K_c the unit-coefficient
  // cell stiffness (Gauss(p+1) points qx/qw on [0,1], S/D the FE_Q(p) basis
  // values/derivatives there, [q*(p+1) + a]); lam, w node-major, nt values.
  void launch_cell_grad(const GridDesc &g, const double *vertices, int nt, const double *lam,
                        const double *w, const double *kappa, const double *qx, const double *qw,
                        const double *S, const double *D, double *grad, hipStream_t s);

  // ------------------------------------------------------- slab error
  // Per-cell {L2^2, Linf, H1semi^2} errors of a slab against the
  // manufactured heat solution: nq Gauss points (qx, qw on [0,1]) with the
  // FE_Q(p) values S and derivatives D there (nq x (p+1), row-major); tpts
  // holds npts rows [time, weight, coefficients of the nt blocks of x, of
  // prev]; cell_err gets 3 values per cell.
  void launch_slab_error(const GridDesc &g, const double *vertices, int nt, const double *x,
                         const double *prev, int nq, const double *qx, const double *qw,
                         const double *S, const double *D, int npts, const double *tpts,
                         double freq, double *cell_err, hipStream_t s);

  // ------------------------------------------------------- wave velocity
  // TimeIntegratorWave velocity recovery per node (node-major u, v with
  // nt = nts * ni values per node): for every step it,
  //   v_it = AixB u_it + AixG pu       (dG,  pu: previous step's last u)
  //   v_it = AixB u_it + AixG pv + AixZ pu   (cG, pv: previous step's last v)
  // with pu/pv = prev_u/prev_v for the first step.  AixB: ni x ni
  // row-major, AixG, AixZ: ni values (device, FP64).
  void launch_wave_velocity(long n_nodes, int nts, int ni, bool dg, const double *AixB,
                            const double *AixG, const double *AixZ, const double *u,
                            const double *prev_u, const double *prev_v, double *v,
                            hipStream_t s);

  // ------------------------------------------------- dense coarse operator
  // M[i + i*n] += a (column-major n x n);  x[0..n) *= a
  template <typename C>
  void launch_add_diagonal(long n, C a, C *M, hipStream_t s);
  template <typename C>
  void launch_scale(long n, C a, C *x, hipStream_t s);
  // Probing of a local linear map (columns of its dense matrix in batches).
  // A colour is (c[0..dim), t): the inputs (node, t) of the grid `in` with
  // node lattice position g = c (mod stride) in every direction.
  // launch_probe_vector: x = the sum of the colour's unit vectors.
  // launch_probe_store: y = the map applied to that x; each output entry i
  // (node of the grid `out` at lattice position o) belongs to the one input
  // g of the colour with |scale_num * g - scale_den * o| <= scale_den *
  // radius in every direction (none if there is no such g): M[j * ld + i] =
  // y[i] for its input column j = node(g) * nt_in + t.  A nonzero y[i] that
  // belongs to no input (supports not disjoint) is counted in *bad.
  template <typename T>
  void launch_probe_vector(const GridDesc &in, int nt_in, const int *c, int t, int stride, T *x,
                           hipStream_t s);
  template <typename T, typename C>
  void launch_probe_store(const GridDesc &in, int nt_in, const GridDesc &out, int nt_out,
                          const int *c, int t, int stride, int scale_num, int scale_den,
                          int radius, const T *y, C *M, long ld, int *bad, hipStream_t s);
  // B (n x m, column-major, leading dimension n) = A^T (A: m x n, leading m)
  template <typename C>
  void launch_transpose(long m, long n, const C *A, C *B, hipStream_t s);
  // R[i*ld + j] = T(M[i + j*n]) for i, j < n (column-major n x n FP64 into
  // row-major precision-T rows of ld >= n entries, padding zeroed)
  template <typename T, typename C>
  void launch_pack_rows(long n, const C *M, T *R, long ld, hipStream_t s);
  // y = R x for the row-major R above (one wavefront per row, FP64 sums);
  // ld must be a multiple of 16 bytes / sizeof(T)
  template <typename T>
  void launch_gemv_rows(long n, long ld, const T *R, const T *x, T *y, hipStream_t s);
} // namespace stfem
