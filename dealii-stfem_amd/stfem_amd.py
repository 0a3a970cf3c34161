"""ctypes binding of libstfem_amd.so (C ABI declared in include/stfem.h).

Python mirror of the reference's host interface for the space-time slab
solve (nlsmrg/dealii-stfem tests/include/{fe_time,operators,gmg,
time_integrators}.h).  Device vectors are torch tensors on the problem's
device; the library does the arithmetic in hand-written HIP kernels.  There
is no CPU fallback: every call fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstfem_amd.so")
_LIB = None

DG = 2
CGP = 1
FINE = -1  # level index of the FP64 fine system

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p

# exported entry points and their argument types
_SIGS = {
    "stfem_version": ([], ctypes.c_int),
    "stfem_last_error": ([], ctypes.c_char_p),
    "stfem_fe_time_weights": ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                               _dp, _dp, _dp, _dp], ctypes.c_int),
    "stfem_dg_weights": ([ctypes.c_int, _dp, _dp, _dp], ctypes.c_int),
    "stfem_cg_weights": ([ctypes.c_int, _dp, _dp], ctypes.c_int),
    "stfem_fe_time_weights_wave": ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                    _dp, _dp, _dp, _dp, _dp], ctypes.c_int),
    "stfem_time_prolongation": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
    "stfem_time_restriction": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
    "stfem_time_projection": ([ctypes.c_int] * 4 + [_dp, _ip, _ip], ctypes.c_int),
    "stfem_quadrature": ([ctypes.c_int, ctypes.c_int, _dp, _dp], ctypes.c_int),
    "stfem_time_mg_sequence": ([ctypes.c_int] * 6 + [_ip, _ip], ctypes.c_int),
    "stfem_problem_create": ([ctypes.c_char_p, _dp, _dp, ctypes.c_int,
                              ctypes.POINTER(_vp)], ctypes.c_int),
    "stfem_problem_destroy": ([_vp], ctypes.c_int),
    "stfem_problem_info": ([_vp, _lp], ctypes.c_int),
    "stfem_problem_tau": ([_vp, _dp], ctypes.c_int),
    "stfem_level_info": ([_vp, ctypes.c_int, _lp], ctypes.c_int),
    "stfem_level_omega": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "stfem_st_vmult": ([_vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "stfem_st_tvmult": ([_vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "stfem_vanka_vmult": ([_vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "stfem_prolongate_and_add": ([_vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "stfem_restrict_and_add": ([_vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "stfem_gmg_vmult": ([_vp, _vp, _vp], ctypes.c_int),
    "stfem_fgmres_solve": ([_vp, _vp, _vp, _ip, _dp, ctypes.c_int], ctypes.c_int),
    "stfem_solve_slab": ([_vp, _vp, _vp, ctypes.c_double, _ip], ctypes.c_int),
    "stfem_solve_slab_wave": ([_vp, _vp, _vp, _vp, _vp, ctypes.c_double, _ip], ctypes.c_int),
    "stfem_slab_error": ([_vp, _vp, _vp, ctypes.c_double, _dp], ctypes.c_int),
    "stfem_run_heat": ([_vp, _dp, _ip, ctypes.c_int], ctypes.c_int),
    "stfem_set_load_vector": ([_vp, _dp], ctypes.c_int),
    "stfem_forward_trajectory": ([_vp, ctypes.c_int, _vp, _ip], ctypes.c_int),
    "stfem_misfit": ([_vp, _vp, ctypes.c_int, _dp, _ip], ctypes.c_int),
    "stfem_adjoint_gradient": ([_vp, _vp, ctypes.c_int, _vp, _dp, _ip], ctypes.c_int),
    "stfem_set_cell_coefficient": ([_vp, _dp], ctypes.c_int),
    "stfem_node_coordinates": ([_vp, _dp], ctypes.c_int),
    "stfem_bytes_model": ([_vp, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
    "stfem_blocks_to_nodes": ([ctypes.c_long, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "stfem_nodes_to_blocks": ([ctypes.c_long, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "stfem_time_kernel": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp],
                          ctypes.c_int),
    "stfem_time_slabs": ([_vp, ctypes.c_int, _dp, _ip], ctypes.c_int),
    "stfem_spec_normalize": ([ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
    "stfem_phase_timings": ([_vp, ctypes.c_int, _dp, _lp], ctypes.c_int),
    # object-level boundary
    "stfem_context_create": ([ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "stfem_context_destroy": ([_vp], ctypes.c_int),
    "stfem_space_hierarchy_create": ([_vp, ctypes.c_char_p, _dp, _dp, ctypes.c_int,
                                      ctypes.POINTER(_vp), _ip], ctypes.c_int),
    "stfem_space_destroy": ([_vp], ctypes.c_int),
    "stfem_space_info": ([_vp, _lp], ctypes.c_int),
    "stfem_space_constrained": ([_vp, ctypes.c_void_p], ctypes.c_int),
    "stfem_stop_create": ([_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int,
                           ctypes.POINTER(_vp)], ctypes.c_int),
    "stfem_stop_destroy": ([_vp], ctypes.c_int),
    "stfem_stop_info": ([_vp, _lp], ctypes.c_int),
    "stfem_stop_vmult": ([_vp, _vp, _vp], ctypes.c_int),
    "stfem_stop_tvmult": ([_vp, _vp, _vp], ctypes.c_int),
    "stfem_stop_vmult_noalloc": ([_vp, _vp, _vp], ctypes.c_int),
    "stfem_stop_vmult_add": ([_vp, _vp, _vp], ctypes.c_int),
    "stfem_vanka_create": ([_vp, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "stfem_vanka_destroy": ([_vp], ctypes.c_int),
    "stfem_vanka_apply": ([_vp, _vp, _vp], ctypes.c_int),
    "stfem_time_transfer_create": ([_vp] + [ctypes.c_int] * 6 + [ctypes.POINTER(_vp)],
                                   ctypes.c_int),
    "stfem_space_transfer_create": ([_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)],
                                    ctypes.c_int),
    "stfem_transfer_destroy": ([_vp], ctypes.c_int),
    "stfem_transfer_prolongate_and_add": ([_vp, _vp, _vp], ctypes.c_int),
    "stfem_transfer_restrict_and_add": ([_vp, _vp, _vp], ctypes.c_int),
    "stfem_gmg_create": ([ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                          ctypes.POINTER(_vp), ctypes.c_char_p, _dp, ctypes.POINTER(_vp)],
                         ctypes.c_int),
    "stfem_gmg_destroy": ([_vp], ctypes.c_int),
    "stfem_gmg_omega": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "stfem_gmg_apply": ([_vp, _vp, _vp], ctypes.c_int),
    "stfem_fgmres_create": ([_vp, ctypes.c_long, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                             ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "stfem_fgmres_destroy": ([_vp], ctypes.c_int),
    "stfem_fgmres_run": ([_vp, _vp, _vp, _vp, _vp, _ip, _dp, ctypes.c_int], ctypes.c_int),
}


class StfemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"stfem error {code}: {msg}")
        self.code = code


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run python dealii-stfem_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise StfemError(rc, load().stfem_last_error().decode())
    return rc


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _ptr(t):
    """Device pointer of a torch tensor (contiguous)."""
    assert t.is_contiguous()
    return ctypes.c_void_p(t.data_ptr())


def nt_intvl(type_, r):
    return r if type_ == CGP else r + 1


# ------------------------------------------------------------------ time FE
def fe_time_weights(type_, r, tau, nts=1):
    n = nt_intvl(type_, r) * nts
    A, B, G, Z = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, 1)), np.zeros((n, 1))
    _check(load().stfem_fe_time_weights(type_, r, tau, nts, _p(A), _p(B), _p(G), _p(Z)))
    return A, B, G, Z


def dg_weights(r):
    m, d, j = np.zeros((r + 1, r + 1)), np.zeros((r + 1, r + 1)), np.zeros((r + 1, 1))
    _check(load().stfem_dg_weights(r, _p(m), _p(d), _p(j)))
    return m, d, j


def cg_weights(r):
    m, d = np.zeros((r, r + 1)), np.zeros((r, r + 1))
    _check(load().stfem_cg_weights(r, _p(m), _p(d)))
    return m, d


def fe_time_weights_wave(type_, r, tau, nts=1):
    n = nt_intvl(type_, r) * nts
    out = [np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, 1)), np.zeros((n, 1)),
           np.zeros((n, 1))]
    _check(load().stfem_fe_time_weights_wave(type_, r, tau, nts, *[_p(o) for o in out]))
    return out


def time_prolongation(type_, r, nts=2):
    ni = nt_intvl(type_, r)
    P = np.zeros((ni * nts, ni * nts // 2))
    _check(load().stfem_time_prolongation(type_, r, nts, _p(P)))
    return P


def time_restriction(type_, r, nts=2):
    ni = nt_intvl(type_, r)
    R = np.zeros((ni * nts // 2, ni * nts))
    _check(load().stfem_time_restriction(type_, r, nts, _p(R)))
    return R


def time_projection(type_, r_src, r_dst, nts=1):
    rows, cols = ctypes.c_int(), ctypes.c_int()
    _check(load().stfem_time_projection(type_, r_src, r_dst, nts, None, ctypes.byref(rows),
                                        ctypes.byref(cols)))
    P = np.zeros((rows.value, cols.value))
    _check(load().stfem_time_projection(type_, r_src, r_dst, nts, _p(P), ctypes.byref(rows),
                                        ctypes.byref(cols)))
    return P


def quadrature(kind, n):
    x, w = np.zeros(n), np.zeros(n)
    _check(load().stfem_quadrature(kind, n, _p(x), _p(w)))
    return x, w


def time_mg_sequence(n_sp, kmax, kmin, nts, nts_min, time_before_space=False):
    n = ctypes.c_int()
    _check(load().stfem_time_mg_sequence(n_sp, kmax, kmin, nts, nts_min, int(time_before_space),
                                         None, ctypes.byref(n)))
    out = (ctypes.c_int * max(n.value, 1))()
    _check(load().stfem_time_mg_sequence(n_sp, kmax, kmin, nts, nts_min, int(time_before_space),
                                         out, ctypes.byref(n)))
    return [out[i] for i in range(n.value)]


# ------------------------------------------------------------------ problem
def spec_string(spec):
    def fmt(v):
        if isinstance(v, (list, tuple)):
            return ",".join(str(x) for x in v)
        return str(v)
    return ";".join(f"{k}={fmt(v)}" for k, v in spec.items())


def spec_normalize(spec):
    """Parameters::parse without building anything: the spec (a dict, a
    "key=value;..." string or the text of a JSON parameter file) with every
    key resolved, as a dict of strings."""
    text = spec if isinstance(spec, str) else spec_string(spec)
    buf = ctypes.create_string_buffer(8192)
    _check(load().stfem_spec_normalize(text.encode(), buf, len(buf)))
    return dict(kv.split("=", 1) for kv in buf.value.decode().split(";"))


class Problem:
    """Hierarchy + solver built from a spec dict (keys listed in include/stfem.h)."""

    def __init__(self, spec, vertices=None, kappa=None, device=0):
        self._v = None if vertices is None else np.ascontiguousarray(vertices, np.float64)
        self._k = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
        h = ctypes.c_void_p()
        text = spec if isinstance(spec, str) else spec_string(spec)
        _check(load().stfem_problem_create(text.encode(), _p(self._v), _p(self._k),
                                           device, ctypes.byref(h)))
        self.h = h
        self.device = device
        self.dim = int(spec_normalize(text)["dim"])
        info = np.zeros(6, np.int64)
        _check(load().stfem_problem_info(self.h, info.ctypes.data_as(_lp)))
        self.n_levels, self.nt, self.n_nodes, self.n_cells, self.level_size, self.nts = \
            [int(x) for x in info]
        tau = ctypes.c_double()
        _check(load().stfem_problem_tau(self.h, ctypes.byref(tau)))
        self.tau = tau.value

    def close(self):
        if getattr(self, "h", None):
            load().stfem_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def level_info(self, l):
        info = np.zeros(12, np.int64)
        _check(load().stfem_level_info(self.h, l, info.ctypes.data_as(_lp)))
        keys = ["n_nodes", "nt", "n_cells", "r", "nts", "mg_type", "nc0", "nc1", "nc2",
                "number_size", "in_dense", "dense_dofs"]
        return dict(zip(keys, [int(x) for x in info]))

    def omega(self, l):
        w = ctypes.c_double()
        _check(load().stfem_level_omega(self.h, l, ctypes.byref(w)))
        return w.value

    def vmult(self, level, dst, src):
        _check(load().stfem_st_vmult(self.h, level, _ptr(dst), _ptr(src)))

    def Tvmult(self, level, dst, src):
        _check(load().stfem_st_tvmult(self.h, level, _ptr(dst), _ptr(src)))

    def vanka(self, level, dst, src):
        _check(load().stfem_vanka_vmult(self.h, level, _ptr(dst), _ptr(src)))

    def prolongate_and_add(self, level, dst, src):
        _check(load().stfem_prolongate_and_add(self.h, level, _ptr(dst), _ptr(src)))

    def restrict_and_add(self, level, dst, src):
        _check(load().stfem_restrict_and_add(self.h, level, _ptr(dst), _ptr(src)))

    def gmg_vmult(self, dst, src):
        _check(load().stfem_gmg_vmult(self.h, _ptr(dst), _ptr(src)))

    def fgmres(self, x, b, hist_cap=256):
        its = ctypes.c_int()
        hist = np.zeros(hist_cap)
        rc = load().stfem_fgmres_solve(self.h, _ptr(x), _ptr(b), ctypes.byref(its), _p(hist),
                                       hist_cap)
        return rc, its.value, hist[: its.value + 1]

    def solve_slab(self, x, prev, time):
        its = ctypes.c_int()
        rc = load().stfem_solve_slab(self.h, _ptr(x), _ptr(prev), float(time), ctypes.byref(its))
        return rc, its.value

    def solve_slab_wave(self, x, v, prev_u, prev_v, time):
        """One wave slab: displacement x and velocity v (nt values per node)."""
        its = ctypes.c_int()
        rc = load().stfem_solve_slab_wave(self.h, _ptr(x), _ptr(v), _ptr(prev_u), _ptr(prev_v),
                                          float(time), ctypes.byref(its))
        return rc, its.value

    def slab_error(self, x, prev, time):
        """{L2^2, Linf, H1semi^2} of one slab against the manufactured solution."""
        out = np.zeros(3)
        _check(load().stfem_slab_error(self.h, _ptr(x), _ptr(prev), float(time), _p(out)))
        return out

    def run_heat(self, max_slabs=4096):
        """The heat time loop with errors (same record as the oracle's run_heat)."""
        stats = np.zeros(7)
        per = (ctypes.c_int * max_slabs)()
        rc = load().stfem_run_heat(self.h, _p(stats), per, max_slabs)
        if rc not in (0, -3, -5):
            _check(rc)
        n = int(stats[0])
        return dict(rc=rc, slabs=n, total_its=int(stats[1]), L2=stats[2], Linf=stats[3],
                    H1=stats[4], ndofs=int(stats[5]), nblocks=int(stats[6]),
                    its=[per[i] for i in range(min(n, max_slabs))])

    def node_coordinates(self):
        """Support points of the fine nodes, (n_nodes, dim) host array."""
        xyz = np.zeros(self.n_nodes * self.dim)
        _check(load().stfem_node_coordinates(self.h, _p(xyz)))
        return xyz.reshape(self.n_nodes, self.dim)

    # ---- adjoint of the heat slab sequence (dG)
    def forward_trajectory(self, n_slabs, traj):
        """Forward slab solutions from zero initial values into the device
        tensor traj (n_slabs * n_nodes * nt); returns iterations per slab."""
        its = np.zeros(n_slabs, np.int32)
        _check(load().stfem_forward_trajectory(self.h, int(n_slabs), _ptr(traj),
                                               its.ctypes.data_as(_ip)))
        return its

    def misfit(self, obs, n_slabs):
        J = ctypes.c_double()
        its = np.zeros(n_slabs, np.int32)
        _check(load().stfem_misfit(self.h, _ptr(obs), int(n_slabs), ctypes.byref(J),
                                   its.ctypes.data_as(_ip)))
        return J.value, its

    def adjoint_gradient(self, obs, n_slabs, grad):
        """dJ/dlog(kappa) into the device tensor grad (n_cells); returns
        (J, per-slab forward and adjoint iterations)."""
        J = ctypes.c_double()
        its = np.zeros(2 * n_slabs, np.int32)
        _check(load().stfem_adjoint_gradient(self.h, _ptr(obs), int(n_slabs), _ptr(grad),
                                             ctypes.byref(J), its.ctypes.data_as(_ip)))
        return J.value, its

    def set_cell_coefficient(self, kappa):
        k = np.ascontiguousarray(kappa, np.float64)
        _check(load().stfem_set_cell_coefficient(self.h, _p(k)))

    def set_load_vector(self, h):
        h = np.ascontiguousarray(h, np.float64)
        _check(load().stfem_set_load_vector(self.h, _p(h)))

    def bytes_model(self, what, level=0):
        b = ctypes.c_double()
        _check(load().stfem_bytes_model(self.h, what, level, ctypes.byref(b)))
        return b.value

    def time_kernel(self, what, level=0, warmup=3, reps=20):
        ms = ctypes.c_double()
        _check(load().stfem_time_kernel(self.h, what, level, warmup, reps, ctypes.byref(ms)))
        return ms.value

    def phase_timings(self, reset=False):
        """{phase: (ms, calls)} for vmult, vanka, gmg, step (spec print_timing=1)."""
        ms = np.zeros(4)
        calls = np.zeros(4, np.int64)
        _check(load().stfem_phase_timings(self.h, int(reset), _p(ms), calls.ctypes.data_as(_lp)))
        return {k: (float(ms[i]), int(calls[i]))
                for i, k in enumerate(("vmult", "vanka", "gmg", "step"))}

    def time_slabs(self, n_slabs, allow_unconverged=False):
        """(ms, iterations) of n_slabs heat slabs; with allow_unconverged,
        (ms, iterations, converged) and slabs stopped by the iteration limit
        are accepted."""
        ms = ctypes.c_double()
        its = ctypes.c_int()
        rc = load().stfem_time_slabs(self.h, n_slabs, ctypes.byref(ms), ctypes.byref(its))
        if allow_unconverged and rc == -3:
            return ms.value, its.value, False
        _check(rc)
        return (ms.value, its.value, True) if allow_unconverged else (ms.value, its.value)


# ------------------------------------------------------------ object boundary
class _Handle:
    """Owner of one library handle; destroyed with the Python object."""
    _destroy = None

    def __init__(self, h, keep=()):
        self.h = h
        self._keep = keep  # Python-side references only; the library keeps its own

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            getattr(load(), self._destroy)(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _new(fn, *args):
    h = ctypes.c_void_p()
    _check(getattr(load(), fn)(*args, ctypes.byref(h)))
    return h


def _mat(a):
    return np.ascontiguousarray(a, np.float64)


class Context(_Handle):
    """Device and stream shared by the objects built on it."""
    _destroy = "stfem_context_destroy"

    def __init__(self, device=0):
        super().__init__(_new("stfem_context_create", device))


class Space(_Handle):
    """One level's DoFHandler: structured mesh, FE_Q(p), coefficient."""
    _destroy = "stfem_space_destroy"

    @staticmethod
    def hierarchy(ctx, spec, vertices=None, kappa=None):
        text = spec if isinstance(spec, str) else spec_string(spec)
        v = None if vertices is None else np.ascontiguousarray(vertices, np.float64)
        k = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
        arr = (ctypes.c_void_p * 16)()
        n = ctypes.c_int()
        _check(load().stfem_space_hierarchy_create(ctx.h, text.encode(), _p(v), _p(k), 16, arr,
                                                   ctypes.byref(n)))
        return [Space(ctypes.c_void_p(arr[i])) for i in range(n.value)]

    def info(self):
        i = np.zeros(8, np.int64)
        _check(load().stfem_space_info(self.h, i.ctypes.data_as(_lp)))
        keys = ["dim", "p", "nc0", "nc1", "nc2", "n_nodes", "n_cells", "dirichlet"]
        return dict(zip(keys, [int(x) for x in i]))

    def constrained(self):
        m = np.zeros(self.info()["n_nodes"], np.uint8)
        _check(load().stfem_space_constrained(self.h, m.ctypes.data_as(ctypes.c_void_p)))
        return m.astype(bool)


class SystemMatrix(_Handle):
    """Alpha (x) K + Beta (x) M on a space (operators.h:71-256)."""
    _destroy = "stfem_stop_destroy"

    def __init__(self, space, alpha, beta, precision=8):
        a, b = _mat(alpha), _mat(beta)
        super().__init__(_new("stfem_stop_create", space.h, a.shape[0], a.shape[1], _p(a), _p(b),
                              precision), keep=(space,))

    def vmult(self, dst, src):
        _check(load().stfem_stop_vmult(self.h, _ptr(dst), _ptr(src)))

    def Tvmult(self, dst, src):
        _check(load().stfem_stop_tvmult(self.h, _ptr(dst), _ptr(src)))

    def vmult_noalloc(self, dst, src):
        """vmult with the library's fixed scratch array (no allocation)."""
        _check(load().stfem_stop_vmult_noalloc(self.h, _ptr(dst), _ptr(src)))

    def vmult_add(self, dst, src):
        _check(load().stfem_stop_vmult_add(self.h, _ptr(dst), _ptr(src)))


class PreconditionVanka(_Handle):
    """Cell-patch Vanka smoother (gmg.h:402-540)."""
    _destroy = "stfem_vanka_destroy"

    def __init__(self, space, alpha, beta, precision=8, factored=True, device_setup=3):
        a, b = _mat(alpha), _mat(beta)
        super().__init__(_new("stfem_vanka_create", space.h, a.shape[0], _p(a), _p(b), precision,
                              int(factored), device_setup), keep=(space,))

    def vmult(self, dst, src):
        _check(load().stfem_vanka_apply(self.h, _ptr(dst), _ptr(src)))


class Transfer(_Handle):
    """Two-level transfer in time (MGTwoLevelTransferST) or space
    (MGTwoLevelBlockTransfer)."""
    _destroy = "stfem_transfer_destroy"

    @staticmethod
    def time(space, type_, r, nts, kind, restrict_is_transpose=True, precision=8):
        return Transfer(_new("stfem_time_transfer_create", space.h, type_, r, nts, kind,
                             int(restrict_is_transpose), precision), keep=(space,))

    @staticmethod
    def space(fine, coarse, nt, precision=8):
        return Transfer(_new("stfem_space_transfer_create", fine.h, coarse.h, nt, precision),
                        keep=(fine, coarse))

    def prolongate_and_add(self, dst, src):
        _check(load().stfem_transfer_prolongate_and_add(self.h, _ptr(dst), _ptr(src)))

    def restrict_and_add(self, dst, src):
        _check(load().stfem_transfer_restrict_and_add(self.h, _ptr(dst), _ptr(src)))


class GMG(_Handle):
    """Space-time multigrid preconditioner (gmg.h:652-906)."""
    _destroy = "stfem_gmg_destroy"

    def __init__(self, ops, smoothers, transfers, options="", omega=None):
        n = len(ops)
        arr = lambda objs: (ctypes.c_void_p * n)(*[o.h.value if o is not None else None
                                                   for o in objs])
        tr = [None] + list(transfers[1:]) if len(transfers) == n else [None] + list(transfers)
        om = None if omega is None else np.ascontiguousarray(omega, np.float64)
        text = options if isinstance(options, str) else spec_string(options)
        super().__init__(_new("stfem_gmg_create", n, arr(ops), arr(smoothers), arr(tr),
                              text.encode(), _p(om)), keep=(ops, smoothers, transfers))
        self.n_levels = n

    def omega(self, l):
        w = ctypes.c_double()
        _check(load().stfem_gmg_omega(self.h, l, ctypes.byref(w)))
        return w.value

    def vmult(self, dst, src):
        _check(load().stfem_gmg_apply(self.h, _ptr(dst), _ptr(src)))


class SolverFGMRES(_Handle):
    """SolverFGMRES with ReductionControl (time_integrators.h:44, :113)."""
    _destroy = "stfem_fgmres_destroy"

    def __init__(self, ctx, n, maxiter=200, abstol=1e-12, reltol=1e-12, basis=30,
                 orthogonalization=1):
        super().__init__(_new("stfem_fgmres_create", ctx.h, int(n), maxiter, abstol, reltol,
                              basis, orthogonalization), keep=(ctx,))

    def solve(self, A, x, b, P=None, hist_cap=256):
        """(status, iterations, residual history); x holds the start vector."""
        its = ctypes.c_int()
        hist = np.zeros(hist_cap)
        rc = load().stfem_fgmres_run(self.h, A.h, P.h if P is not None else None, _ptr(x),
                                     _ptr(b), ctypes.byref(its), _p(hist), hist_cap)
        if rc not in (0, -3, -5):  # argument errors raise; solver outcomes are returned
            _check(rc)
        return rc, its.value, hist[: its.value + 1]
