"""stfem_stop_vmult_noalloc (the operator application that allocates
nothing: its cell-major intermediate results live in a fixed device array of
the library) against stfem_stop_vmult on the same operators: equal to the
last bit, on two calls in a row, the second with a smaller operator (the
array keeps the first call's contents), in FP64 and FP32; operators whose
intermediate results exceed the array are refused with a message."""
import numpy as np
import pytest

from helpers import box_vertices, log_uniform

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import stfem_amd as st  # noqa: E402

DEV = "cuda:0"


def op_on(ctx, refine, p, nt, prec, seed):
    nc = [1 << refine] * 3
    X = box_vertices(3, nc, perturb=0.2, seed=seed)
    kap = log_uniform(int(np.prod(nc)), 1e-2, 1e2, seed=seed + 1)
    spaces = st.Space.hierarchy(ctx, dict(dim=3, refine=refine, p=p), X, kap)
    A, B, _, _ = st.fe_time_weights(st.DG, nt - 1, 0.1, 1)
    return st.SystemMatrix(spaces[-1], A, B, prec), spaces[-1].info()["n_nodes"] * nt


@pytest.mark.parametrize("prec", [8, 4])
def test_noalloc_vmult_matches_vmult(prec):
    ctx = st.Context()
    dt = torch.float64 if prec == 8 else torch.float32
    rng = np.random.default_rng(4)
    # a larger operator first (32^3 Q3 dG(2): 50 MB of FP64 cell results), then
    # a smaller one (4^3 Q2 dG(1)) over the leftovers of the first call
    for refine, p, nt in [(5, 3, 3), (2, 2, 2)]:
        op, n = op_on(ctx, refine, p, nt, prec, seed=refine)
        x = torch.tensor(rng.standard_normal(n), dtype=dt, device=DEV)
        y_ref, y = torch.zeros_like(x), torch.full_like(x, float("nan"))
        print(f"[noalloc prec={prec}] refine {refine} Q{p} nt {nt}: vmult", flush=True)
        op.vmult(y_ref, x)
        print(f"[noalloc prec={prec}] refine {refine} Q{p} nt {nt}: vmult_noalloc", flush=True)
        op.vmult_noalloc(y, x)
        assert torch.equal(y, y_ref), (y - y_ref).abs().max().item()
        op.close()


def test_noalloc_refuses_oversized_operators():
    ctx = st.Context()
    # 32^3 Q4 dG(3) in FP64: 32768 * 125 * 4 * 8 bytes = 131 MB > 64 MiB
    op, n = op_on(ctx, 5, 4, 4, 8, seed=1)
    x = torch.zeros(n, dtype=torch.float64, device=DEV)
    y = torch.zeros_like(x)
    print("[noalloc] oversized operator", flush=True)
    with pytest.raises(st.StfemError) as e:
        op.vmult_noalloc(y, x)
    assert e.value.code == -1 and "exceeds the fixed scratch" in str(e.value)
    op.close()
