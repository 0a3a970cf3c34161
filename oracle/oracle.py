"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  The oracle restates nlsmrg/dealii-stfem's algorithm on
the CPU (oracle/stfem_oracle.cpp cites the reference file:line for each part).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

DG = 2
CGP = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def lib_path():
    return os.path.join(_HERE, "build", "liboracle.so")


def load():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.oracle_problem_create.restype = ctypes.c_void_p
        L.oracle_problem_create.argtypes = [ctypes.c_char_p, _dp, _dp]
        L.oracle_problem_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_n_levels.argtypes = [ctypes.c_void_p]
        L.oracle_tau.restype = ctypes.c_double
        L.oracle_tau.argtypes = [ctypes.c_void_p]
        L.oracle_level_info.argtypes = [ctypes.c_void_p, ctypes.c_int, _lp]
        L.oracle_level_omega.restype = ctypes.c_double
        L.oracle_level_omega.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_level_time_matrices.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.oracle_fine_vmult.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.oracle_level_apply.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.oracle_gmg_vmult.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_fine_matrix.restype = ctypes.c_long
        L.oracle_fine_matrix.argtypes = [ctypes.c_void_p, ctypes.c_int, _lp, _lp, _dp]
        L.oracle_level_matrix.restype = ctypes.c_long
        L.oracle_level_matrix.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _lp, _lp,
                                          _dp]
        L.oracle_level_lambda.restype = ctypes.c_double
        L.oracle_level_lambda.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_fgmres.argtypes = [ctypes.c_void_p, _dp, _dp, _ip, _dp, ctypes.c_int]
        L.oracle_run_heat.argtypes = [ctypes.c_void_p, _dp, _ip, ctypes.c_int]
        L.oracle_fe_time_weights.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_int, _dp, _dp, _dp, _dp]
        L.oracle_dg_weights.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.oracle_cg_weights.argtypes = [ctypes.c_int, _dp, _dp]
        L.oracle_fe_time_weights_wave.argtypes = [ctypes.c_int] * 3 + [_dp] * 5
        L.oracle_wave_from.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp,
                                       ctypes.c_int] + [_dp] * 5
        L.oracle_time_prolongation.argtypes = [ctypes.c_int] * 3 + [_dp]
        L.oracle_time_restriction.argtypes = [ctypes.c_int] * 3 + [_dp]
        L.oracle_time_projection.argtypes = [ctypes.c_int] * 4 + [_dp, _ip, _ip]
        L.oracle_quadrature.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.oracle_mg_sequence.argtypes = [ctypes.c_int] * 6 + [_ip]
        _LIB = L
    return _LIB


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def num_threads():
    """Threads of the oracle's OpenMP loops (spmv, Vanka setup and apply)."""
    L = load()
    L.oracle_num_threads.restype = ctypes.c_int
    return L.oracle_num_threads()


def nt_intvl(type_, r):
    return r if type_ == CGP else r + 1


def fe_time_weights(type_, r, tau, nts):
    n = nt_intvl(type_, r) * nts
    A, B = np.zeros((n, n)), np.zeros((n, n))
    G, Z = np.zeros((n, 1)), np.zeros((n, 1))
    load().oracle_fe_time_weights(type_, r, tau, nts, _p(A), _p(B), _p(G), _p(Z))
    return A, B, G, Z


def dg_weights(r):
    n = r + 1
    m, d, j = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, 1))
    load().oracle_dg_weights(r, _p(m), _p(d), _p(j))
    return m, d, j


def cg_weights(r):
    m, d = np.zeros((r, r + 1)), np.zeros((r, r + 1))
    load().oracle_cg_weights(r, _p(m), _p(d))
    return m, d


def wave_weights(type_, r, nts):
    n = nt_intvl(type_, r) * nts
    out = [np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, 1)), np.zeros((n, 1)),
           np.zeros((n, 1))]
    load().oracle_fe_time_weights_wave(type_, r, nts, *[_p(o) for o in out])
    return out


def wave_from(type_, A, B, G, Z, nts=1):
    m = A.shape[0]
    n = m * nts
    out = [np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, 1)), np.zeros((n, 1)),
           np.zeros((n, 1))]
    A, B, G = [np.ascontiguousarray(x, dtype=np.float64) for x in (A, B, G)]
    Z = None if Z is None else np.ascontiguousarray(Z, dtype=np.float64)
    load().oracle_wave_from(type_, m, _p(A), _p(B), _p(G), _p(Z), nts, *[_p(o) for o in out])
    return out


def time_prolongation(type_, r, nts=2):
    ni = nt_intvl(type_, r)
    P = np.zeros((ni * nts, ni * nts // 2))
    load().oracle_time_prolongation(type_, r, nts, _p(P))
    return P


def time_restriction(type_, r, nts=2):
    ni = nt_intvl(type_, r)
    R = np.zeros((ni * nts // 2, ni * nts))
    load().oracle_time_restriction(type_, r, nts, _p(R))
    return R


def time_projection(type_, r_src, r_dst, nts=1):
    rows, cols = ctypes.c_int(), ctypes.c_int()
    load().oracle_time_projection(type_, r_src, r_dst, nts, None, ctypes.byref(rows),
                                  ctypes.byref(cols))
    P = np.zeros((rows.value, cols.value))
    load().oracle_time_projection(type_, r_src, r_dst, nts, _p(P), ctypes.byref(rows),
                                  ctypes.byref(cols))
    return P


def quadrature(kind, n):
    """kind 0 Gauss, 1 Gauss-Lobatto, 2 Gauss-Radau (right end point)."""
    p, w = np.zeros(n), np.zeros(n)
    load().oracle_quadrature(kind, n, _p(p), _p(w))
    return p, w


def mg_sequence(n_sp, kmax, kmin, nts, nts_min, time_before_space=False):
    n = load().oracle_mg_sequence(n_sp, kmax, kmin, nts, nts_min, int(time_before_space), None)
    out = (ctypes.c_int * max(n, 1))()
    load().oracle_mg_sequence(n_sp, kmax, kmin, nts, nts_min, int(time_before_space), out)
    return [out[i] for i in range(n)]


class Problem:
    """Space-time hierarchy + FGMRES/GMG solver of the oracle.

    spec: dict of keys understood by stfem_oracle.cpp parse_spec().
    """

    def __init__(self, spec, coords=None, kappa=None):
        s = ";".join(f"{k}={v}" for k, v in spec.items())
        self._coords = None if coords is None else np.ascontiguousarray(coords, np.float64)
        self._kappa = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
        self.h = load().oracle_problem_create(s.encode(), _p(self._coords), _p(self._kappa))

    def __del__(self):
        if getattr(self, "h", None):
            load().oracle_problem_destroy(self.h)
            self.h = None

    @property
    def n_levels(self):
        return load().oracle_n_levels(self.h)

    @property
    def tau(self):
        return load().oracle_tau(self.h)

    def level_info(self, l):
        info = np.zeros(9, dtype=np.int64)
        load().oracle_level_info(self.h, l, info.ctypes.data_as(_lp))
        keys = ["n_nodes", "nt", "n_cells", "r", "nts", "mg_type", "nc0", "nc1", "nc2"]
        return dict(zip(keys, [int(v) for v in info]))

    def omega(self, l):
        return load().oracle_level_omega(self.h, l)

    def level_time_matrices(self, l):
        nt = self.level_info(l)["nt"]
        A, B = np.zeros((nt, nt)), np.zeros((nt, nt))
        load().oracle_level_time_matrices(self.h, l, _p(A), _p(B))
        return A, B

    def fine_vmult(self, src, transpose=False):
        src = np.ascontiguousarray(src, np.float64)
        dst = np.zeros_like(src)
        load().oracle_fine_vmult(self.h, _p(dst), _p(src), int(transpose))
        return dst

    def level_apply(self, l, what, src, dst=None):
        """what: 0 vmult, 1 vanka, 2 prolongate_and_add (l-1 -> l), 3 restrict_and_add."""
        src = np.ascontiguousarray(src, np.float64)
        info = self.level_info(l)
        if what in (0, 1):
            dst = np.zeros(info["n_nodes"] * info["nt"])
        elif dst is None:
            if what == 2:
                dst = np.zeros(info["n_nodes"] * info["nt"])
            else:
                ic = self.level_info(l - 1)
                dst = np.zeros(ic["n_nodes"] * ic["nt"])
        dst = np.ascontiguousarray(dst, np.float64)
        load().oracle_level_apply(self.h, l, what, _p(dst), _p(src))
        return dst

    def gmg_vmult(self, src):
        src = np.ascontiguousarray(src, np.float64)
        dst = np.zeros_like(src)
        load().oracle_gmg_vmult(self.h, _p(dst), _p(src))
        return dst

    def fine_matrix(self, which):
        """Assembled fine K (which=0) or M (1) in CSR (deal.II constraint convention)."""
        L = load()
        nnz = L.oracle_fine_matrix(self.h, which, None, None, None)
        n = self.level_info(self.n_levels - 1)["n_nodes"]
        rp = np.zeros(n + 1, np.int64)
        ci = np.zeros(nnz, np.int64)
        v = np.zeros(nnz)
        L.oracle_fine_matrix(self.h, which, rp.ctypes.data_as(_lp), ci.ctypes.data_as(_lp), _p(v))
        return rp, ci, v

    def level_matrix(self, l, which):
        """Assembled K (which=0) or M (1) of MG level l as a dense array
        (deal.II constraint convention: constrained rows/columns reduced to
        the diagonal)."""
        L = load()
        nnz = L.oracle_level_matrix(self.h, l, which, None, None, None)
        n = self.level_info(l)["n_nodes"]
        rp = np.zeros(n + 1, np.int64)
        ci = np.zeros(nnz, np.int64)
        v = np.zeros(nnz)
        L.oracle_level_matrix(self.h, l, which, rp.ctypes.data_as(_lp), ci.ctypes.data_as(_lp),
                              _p(v))
        A = np.zeros((n, n))
        for i in range(n):
            A[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
        return A

    def load_vector(self, time=0.0):
        """int f(time, x) phi_i over the fine level (zero at Dirichlet nodes)."""
        L = load()
        L.oracle_load_vector.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
        out = np.zeros(self.level_info(self.n_levels - 1)["n_nodes"])
        L.oracle_load_vector(self.h, float(time), _p(out))
        return out

    def level_lambda(self, l):
        """Raw power-iteration estimate of the largest eigenvalue of the
        Vanka-preconditioned level operator (0 where none was made)."""
        return load().oracle_level_lambda(self.h, l)

    def fgmres(self, b, x0):
        x = np.ascontiguousarray(x0, np.float64).copy()
        b = np.ascontiguousarray(b, np.float64)
        its = ctypes.c_int()
        hist = np.zeros(256)
        rc = load().oracle_fgmres(self.h, _p(x), _p(b), ctypes.byref(its), _p(hist), 256)
        return rc, x, its.value, hist[: its.value + 1]

    def run_heat(self, max_slabs=4096):
        stats = np.zeros(7)
        per = (ctypes.c_int * max_slabs)()
        rc = load().oracle_run_heat(self.h, _p(stats), per, max_slabs)
        n = int(stats[0])
        return dict(rc=rc, slabs=n, total_its=int(stats[1]), L2=stats[2], Linf=stats[3],
                    H1=stats[4], ndofs=int(stats[5]), nblocks=int(stats[6]),
                    its=[per[i] for i in range(min(n, max_slabs))])


def _solve_slab_sig():
    L = load()
    L.oracle_solve_slab.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, _ip]
    return L


def solve_slab_wave(problem, x, prev, prev_v, time):
    """One wave slab of the oracle; returns (rc, u_new, v_new, iterations)."""
    L = load()
    L.oracle_solve_slab_wave.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, ctypes.c_double,
                                         _ip]
    x = np.ascontiguousarray(x, np.float64).copy()
    v = np.zeros_like(x)
    prev = np.ascontiguousarray(prev, np.float64)
    prev_v = np.ascontiguousarray(prev_v, np.float64)
    its = ctypes.c_int()
    rc = L.oracle_solve_slab_wave(problem.h, _p(x), _p(v), _p(prev), _p(prev_v), float(time),
                                  ctypes.byref(its))
    return rc, x, v, its.value


def slab_error(problem, x, prev, time):
    """{L2^2, Linf, H1semi^2} of one slab (x: nt blocks, prev: one block)."""
    L = load()
    L.oracle_slab_error.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, _dp]
    x = np.ascontiguousarray(x, np.float64)
    prev = np.ascontiguousarray(prev, np.float64)
    out = np.zeros(3)
    L.oracle_slab_error(problem.h, _p(x), _p(prev), float(time), _p(out))
    return out


def solve_slab(problem, x, prev, time):
    """One heat slab of the oracle; returns (rc, x_new, iterations)."""
    x = np.ascontiguousarray(x, np.float64).copy()
    prev = np.ascontiguousarray(prev, np.float64)
    its = ctypes.c_int()
    rc = _solve_slab_sig().oracle_solve_slab(problem.h, _p(x), _p(prev), float(time),
                                             ctypes.byref(its))
    return rc, x, its.value


def forward_trajectory(problem, n_slabs):
    """Forward slab solutions from zero initial values (block-major per slab,
    slabs one after the other); FP64 problems only."""
    info = problem.level_info(problem.n_levels - 1)
    m = info["n_nodes"] * info["nt"]
    out = np.zeros(n_slabs * m)
    L = load()
    L.oracle_forward_trajectory.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    rc = L.oracle_forward_trajectory(problem.h, int(n_slabs), _p(out))
    if rc != 0:
        raise RuntimeError(f"oracle_forward_trajectory: {rc}")
    return out


def misfit(problem, obs, n_slabs):
    L = load()
    L.oracle_misfit.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp]
    obs = np.ascontiguousarray(obs, np.float64)
    J = np.zeros(1)
    rc = L.oracle_misfit(problem.h, _p(obs), int(n_slabs), _p(J))
    if rc != 0:
        raise RuntimeError(f"oracle_misfit: {rc}")
    return float(J[0])


def adjoint_gradient(problem, problem_t, obs, n_slabs):
    """dJ/dlog(kappa) per fine cell and J: problem_t is the same problem built
    with transpose_time=1 (its FGMRES solves the adjoint slab systems)."""
    L = load()
    L.oracle_adjoint_gradient.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _dp, ctypes.c_int,
                                          _dp, _dp]
    obs = np.ascontiguousarray(obs, np.float64)
    info = problem.level_info(problem.n_levels - 1)
    grad = np.zeros(info["n_cells"])
    J = np.zeros(1)
    rc = L.oracle_adjoint_gradient(problem.h, problem_t.h, _p(obs), int(n_slabs), _p(grad), _p(J))
    if rc != 0:
        raise RuntimeError(f"oracle_adjoint_gradient: {rc}")
    return grad, float(J[0])
