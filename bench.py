#!/usr/bin/env python3
"""Benchmark of the MI355X space-time multigrid slab solve.

Metric: space-time DoF/s per slab solve to a 1e-8 relative FP64 residual
with the FP32 multigrid preconditioner inside FP64 FGMRES (value), next to
the plain FP64 solve, plus the FP64 space-time vmult in HBM GB/s.

Workload (default, --workload headline): 3D heat on the unit cube, 64^3
cells with vertices perturbed by up to 0.1h (seeded), Q4 space, dG(3) time
(4 time dofs per slab, 67.9M space-time dofs per slab), T = 1 in 4 slabs,
cellwise log-uniform coefficient in [1e-3, 1e3] (seeded), seeded
Gaussian-bump source.  A step is one slab solve of that run (rhs assembly,
extrapolation, FGMRES + GMG); the run restarts from t = 0 and a zero state
after its 4th slab, so K steps are K slabs of T = 1 runs.

Other workloads: --workload q3 (32^3 cells, Q3/dG(2), 0.2h, [1e-2, 1e2],
T = 1 in 8 slabs) and --workload wave (3D wave, Q2/cG(2)).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dealii-stfem_amd"))

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6     # MI355X dense FP64 (vector and matrix)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def box_vertices(dim, nc, perturb, seed):
    import numpy as np
    n = [c + 1 for c in nc]
    axes = [np.linspace(0.0, 1.0, n[d]) for d in range(dim)]
    grids = np.meshgrid(*axes[::-1], indexing="ij")[::-1]
    X = np.stack([g.reshape(-1) for g in grids], axis=1)
    rng = np.random.default_rng(seed)
    idx = np.indices(n[::-1]).reshape(dim, -1)[::-1]
    interior = np.ones(X.shape[0], dtype=bool)
    for d in range(dim):
        interior &= (idx[d] > 0) & (idx[d] < n[d] - 1)
    for d in range(dim):
        X[interior, d] += perturb * (1.0 / nc[d]) * rng.uniform(-1, 1, interior.sum())
    return np.ascontiguousarray(X)


def workload_q3(refine=5):
    """3D heat on 2^refine cells per side: Q3 space, dG(2) time, vertices
    perturbed by up to 0.2h, cellwise log-uniform coefficient in [1e-2, 1e2],
    seeded Gaussian bumps, T = 1 in 8 slabs; two pre- and two post-smoothing
    steps with the factored Vanka smoother."""
    import numpy as np
    dim = 3
    nc = [1 << refine] * dim
    X = box_vertices(dim, nc, 0.2, seed=2024)
    rng = np.random.default_rng(2025)
    kappa = np.exp(rng.uniform(np.log(1e-2), np.log(1e2), int(np.prod(nc))))
    spec = dict(dim=3, refine=refine, p=3, type=2, r=2, nts=1, r_min=0, tau=0.125,
                end_time=1.0, source=3, n_bumps=4, seed=7, reltol=1e-8, abstol=1e-300,
                maxiter=200, basis=30, vanka_factored=1, smoothing_steps=2)
    desc = (f"3D heat, {1 << refine}^3 cells perturbed 0.2h, Q3 space, dG(2) time, "
            f"log-uniform kappa [1e-2,1e2], Gaussian bumps, T=1 in 8 slabs, tol 1e-8")
    return spec, X, kappa, desc


def workload_headline(refine=6):
    """The large 3D heat workload: 2^refine (64) cells per side perturbed by
    up to 0.1h, Q4 space, dG(3) time (4 time dofs per slab, 67.9M space-time
    dofs per slab at 64^3), T = 1 in 4 slabs, cellwise log-uniform
    coefficient in [1e-3, 1e3], seeded Gaussian bumps.  Three pre- and three
    post-smoothing steps (35 iterations per slab instead of 51 with two, 6 %
    less time); the coarse levels take the arithmetic mean of the finest-level
    coefficient over each quadrature point's share of the cell (with the
    point-sampled coefficient the multigrid does not converge for this
    contrast); the V-cycle from the 4^3-cell level down (19,652 DoFs) is one
    dense operator; an FGMRES basis of 60 vectors (no restart within a slab)."""
    import numpy as np
    dim = 3
    nc = [1 << refine] * dim
    X = box_vertices(dim, nc, 0.1, seed=2026)
    rng = np.random.default_rng(2027)
    kappa = np.exp(rng.uniform(np.log(1e-3), np.log(1e3), int(np.prod(nc))))
    spec = dict(dim=3, refine=refine, p=4, type=2, r=3, nts=1, tau=0.25, end_time=1.0,
                source=3, n_bumps=4, seed=7, reltol=1e-8, abstol=1e-300, maxiter=200,
                basis=60, vanka_factored=1, smoothing_steps=3, coarse_coef=3,
                dense_coarse_dofs=20000)
    desc = (f"3D heat, {1 << refine}^3 cells perturbed 0.1h, Q4 space, dG(3) time, "
            f"log-uniform kappa [1e-3,1e3], Gaussian bumps, T=1 in 4 slabs, tol 1e-8")
    return spec, X, kappa, desc


def workload_wave(refine=5):
    """3D acoustic wave u_tt - div(c^2 grad u) = 0 as the first-order-in-time
    system of tp_01.cc:167-197: unit cube, 2^refine cells per side, Q2 space,
    cG(2) time, tau = 1/8 (8 slabs to T = 1), c^2 = 1 for z < 0.5 and 4 above
    (per cell), Gaussian pulse exp(-|x - x0|^2 / (2 * 0.05^2)) at a seeded
    centre x0 in [0.3, 0.7]^3 as initial displacement, zero initial velocity,
    no source; two pre- and two post-smoothing steps."""
    import numpy as np
    dim = 3
    nc = [1 << refine] * dim
    X = box_vertices(dim, nc, 0.0, seed=0)
    zc = (np.arange(nc[2]) + 0.5) / nc[2]
    kappa = np.repeat(np.where(zc < 0.5, 1.0, 4.0), nc[0] * nc[1])
    centre = np.random.default_rng(2028).uniform(0.3, 0.7, dim)
    spec = dict(dim=3, refine=refine, p=2, type=1, r=2, nts=1, r_min=1, tau=0.125,
                end_time=1.0, problem=1, source=1, reltol=1e-8, abstol=1e-300, maxiter=200,
                basis=30, vanka_factored=1, smoothing_steps=2)
    return spec, X, kappa, centre


def wave_initial(g, centre, width=0.05):
    """Nodal interpolant of the Gaussian pulse (zero on the Dirichlet boundary)."""
    import numpy as np
    xyz = g.node_coordinates()
    u0 = np.exp(-np.sum((xyz - centre) ** 2, axis=1) / (2 * width * width))
    on_bnd = np.any((xyz < 1e-12) | (xyz > 1 - 1e-12), axis=1)
    u0[on_bnd] = 0.0
    return u0


def run_wave(spec, X, kappa, centre, steps, warmup=1):
    """steps wave slabs from the pulse after `warmup` untimed solves of the
    first slab; returns (setup s, ms, its, converged, final u, v, dofs)."""
    import torch
    import stfem_amd as st
    t0 = time.perf_counter()
    g = st.Problem(spec, vertices=X, kappa=kappa)
    setup_s = time.perf_counter() - t0
    n, nt = g.n_nodes, g.nt
    dev = "cuda:0"
    u = torch.zeros(n * nt, dtype=torch.float64, device=dev)
    v = torch.zeros_like(u)
    u0 = torch.tensor(wave_initial(g, centre), dtype=torch.float64, device=dev)
    pu, pv = u0.clone(), torch.zeros(n, dtype=torch.float64, device=dev)
    for _ in range(warmup):  # the loop body once, device code loaded outside the timing
        g.solve_slab_wave(u, v, pu, pv, 0.0)
        pu.copy_(u.view(n, nt)[:, nt - 1])
        pv.copy_(v.view(n, nt)[:, nt - 1])
    pu.copy_(u0)
    pv.zero_()
    its, ok = 0, True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        rc, it = g.solve_slab_wave(u, v, pu, pv, k * g.tau)
        its += it
        ok = ok and rc == 0
        pu.copy_(u.view(n, nt)[:, nt - 1])
        pv.copy_(v.view(n, nt)[:, nt - 1])
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    log(f"wave mg_float={spec.get('mg_float')}: setup {setup_s:.1f} s, {steps} slabs "
        f"{ms:.1f} ms, {its} FGMRES iterations{'' if ok else ' (not converged)'}")
    out = (setup_s, ms, its, ok, u.clone(), v.clone(), n * nt)
    g.close()
    return out


def main_wave(a):
    """The 3D wave workload: FP32 multigrid against FP64, solution difference."""
    import torch
    spec, X, kappa, centre = workload_wave(a.refine or 5)
    for kv in a.set:
        k, v = kv.split("=", 1)
        spec[k] = int(v) if v.lstrip("-").isdigit() else float(v)
    s32, ms32, its32, ok32, u32, v32, nst = run_wave(dict(spec, mg_float=1), X, kappa, centre,
                                                    a.steps, a.warmup)
    s64, ms64, its64, ok64, u64, v64, _ = run_wave(dict(spec, mg_float=0), X, kappa, centre,
                                                  a.steps, a.warmup)
    du = (torch.linalg.norm(u32 - u64) / torch.linalg.norm(u64)).item()
    dv = (torch.linalg.norm(v32 - v64) / torch.linalg.norm(v64)).item()
    out = {
        "metric": "space-time DoF/s per solve to 1e-8 (FP32-MG vs FP64), 3D wave",
        "value": a.steps * nst / (ms32 * 1e-3),
        "unit": "space-time DoF/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms32 / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (FGMRES) + f32 (multigrid)", "data": "synthetic",
        "config": {"workload": (f"3D wave, {1 << (a.refine or 5)}^3 cells, Q2 space, cG(2) "
                                f"time, c^2 = 1 | 4 split at z = 0.5, Gaussian pulse width "
                                f"0.05, v0 = 0, tol 1e-8"),
                   "space_time_dofs_per_slab": nst},
        "fp64_dofs_per_s": a.steps * nst / (ms64 * 1e-3),
        "speedup_fp32mg_vs_fp64": ms64 / ms32,
        "fgmres_its_fp32mg": its32, "fgmres_its_fp64": its64,
        "converged": {"fp32mg": ok32, "fp64": ok64},
        "setup_s": {"fp32mg": s32, "fp64": s64},
        "fp32mg_vs_fp64_rel_l2": {"u": du, "v": dv},
        "fp32mg_vs_fp64_note": ("GPU FP32-multigrid solve vs GPU FP64 solve; the FP64 oracle "
                                "cannot solve this size, its wave slab solves are compared with "
                                "the GPU's (1e-12) at small sizes in the GPU tests"),
    }
    print(json.dumps(out))


def run_gpu(spec, X, kappa, steps, warmup):
    import stfem_amd as st
    t0 = time.perf_counter()
    g = st.Problem(spec, vertices=X, kappa=kappa)
    setup_s = time.perf_counter() - t0
    log(f"setup mg_float={spec.get('mg_float')} {setup_s:.1f} s, {g.n_levels} levels, "
        f"{g.n_nodes * g.nt} space-time dofs")
    if warmup > 0:
        g.time_slabs(warmup, allow_unconverged=True)
    ms, its, ok = g.time_slabs(steps, allow_unconverged=True)
    log(f"timed {steps} slabs: {ms:.1f} ms, {its} FGMRES iterations"
        f"{'' if ok else ' (iteration limit reached)'}")
    return g, setup_s, ms, its, ok


def final_state(g, n_slabs=2):
    """Space-time solution of the first n_slabs slabs from a zero initial
    state (untimed; for the FP32-MG / FP64 comparison)."""
    import torch
    n, nt = g.n_nodes, g.nt
    x = torch.zeros(n * nt, dtype=torch.float64, device="cuda:0")
    prev = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    for k in range(n_slabs):
        prev.copy_(x.view(n, nt)[:, nt - 1])
        g.solve_slab(x, prev, k * g.spec_nts * g.tau)
    torch.cuda.synchronize()
    return x


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(workload, budget_s=30.0, min_s=10.0):
    """The oracle (FP64 CPU restatement of the same algorithm: assembled
    matrices, explicit Vanka inverses, OpenMP over rows and cells on every
    host thread it is given) on a bounded sample of the same discretisation
    and source: the workload's spec on 8^3 cells, the
    slabs of one T = 1 run (slab solves until the budget is used)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc
    refine = 3
    spec, X, kappa, _ = (workload_headline if workload == "headline" else workload_q3)(refine)
    ospec = {k: v for k, v in spec.items()
             if k not in ("source", "vanka_factored", "dense_coarse_dofs")}
    ospec["rhs"] = spec["source"]
    t0 = time.perf_counter()
    o = orc.Problem(ospec, coords=X, kappa=kappa)
    setup = time.perf_counter() - t0
    info = o.level_info(o.n_levels - 1)
    n, nt = info["n_nodes"], info["nt"]
    period = int(round(spec["end_time"] / (spec["nts"] * spec["tau"])))
    x = np.zeros(n * nt)
    t0 = time.perf_counter()
    slabs, its = 0, 0
    # whole T = 1 runs (restarting from zero) until min_s has passed, within
    # the budget
    while slabs % period or time.perf_counter() - t0 < min_s:
        if time.perf_counter() - t0 > budget_s:
            break
        if slabs % period == 0:
            x = np.zeros(n * nt)
        prev = x[(nt - 1) * n:].copy()
        rc, x, k = orc.solve_slab(o, x, prev, (slabs % period) * spec["nts"] * spec["tau"])
        its += k
        slabs += 1
    dt = time.perf_counter() - t0
    # the space-time vmult of the same sample (assembled CSR matrices)
    rng = np.random.default_rng(1)
    xb = rng.standard_normal(n * nt)
    reps, tv0 = 0, time.perf_counter()
    while time.perf_counter() - tv0 < 2.0:
        o.fine_vmult(xb)
        reps += 1
    tv = (time.perf_counter() - tv0) / reps
    thr = orc.num_threads()
    return dict(value=slabs * n * nt / dt, unit="space-time DoF/s", cores=thr, kind="port",
                cpu_model=cpu_model(), host_cpus=os.cpu_count(),
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                sample=(f"oracle FP64 slab solves of the same discretisation and source "
                        f"(spec of the workload on {1 << refine}^3 cells, {n * nt} space-time "
                        f"dofs/slab): {slabs} slabs of T=1 runs ({period} slabs each), {its} "
                        f"FGMRES its, "
                        f"{dt:.1f} s on {thr} threads (setup {setup:.1f} s untimed)"),
                vmult={"value": n * nt / tv, "unit": "space-time DoF/s", "ms": tv * 1e3,
                       "sample": f"oracle FP64 space-time vmult on the same sample, {reps} "
                                 f"applications"})


COPY_BYTES = 512 << 20  # calibration copy of the counter passes


def _counter_csv(d):
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith("counter_collection.csv"):
                return os.path.join(root, f)
    return None


def pmc_traffic(kernel, name, workload, refine, timeout_s):
    """HBM bytes per launch of one kernel from two rocprofv3 counter passes
    (FETCH_SIZE, then WRITE_SIZE) over `bench.py --kernel-only`.  Each counter
    is scaled by what it reads for the calibration copy of COPY_BYTES (the
    first three copy dispatches of the process, before any setup); the
    finest-level launches (largest grid) are the ones reported."""
    import csv
    import re
    import shutil
    import statistics
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not on PATH"
    env = dict(os.environ, TMPDIR="/tmp")
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="stfem_pmc_", dir="/tmp")
        # counters only for the measured kernel and the calibration copy
        cmd = [exe, "--kernel-include-regex", f"copyBuffer|{name}", "--pmc", counter,
               "--output-format", "csv", "-d", d, "-o", "run", "--",
               sys.executable, os.path.abspath(__file__), "--kernel-only", kernel,
               "--workload", workload] + (["--refine", str(refine)] if refine else [])
        log(f"counter pass {counter}: {' '.join(cmd)}")
        try:
            r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                               stderr=subprocess.PIPE, timeout=timeout_s)
        except subprocess.TimeoutExpired:
            return None, f"{counter} pass timed out"
        path = _counter_csv(d)
        if path is None:
            return None, f"{counter} pass wrote no counters (rc {r.returncode})"
        rows = [x for x in csv.DictReader(open(path)) if x["Counter_Name"] == counter]
        shutil.rmtree(d, ignore_errors=True)
        copies = sorted((x for x in rows if "copyBuffer" in x["Kernel_Name"]),
                        key=lambda x: int(x["Dispatch_Id"]))[:3]
        mine = [x for x in rows if re.search(name, x["Kernel_Name"])]
        if len(copies) < 3 or not mine:
            return None, f"{counter}: calibration copy or kernel missing from the counters"
        gmax = max(int(x["Grid_Size"]) for x in mine)
        vals = [float(x["Counter_Value"]) for x in mine if int(x["Grid_Size"]) == gmax]
        scale = COPY_BYTES / (statistics.median(float(x["Counter_Value"]) for x in copies)
                              * 1024.0)
        res[counter] = (statistics.median(vals) * 1024.0 * scale, scale, len(vals))
    fetch, fs, nf = res["FETCH_SIZE"]
    write, ws, nw = res["WRITE_SIZE"]
    return dict(bytes=fetch + write, fetch=fetch, write=write, fetch_scale=fs,
                write_scale=ws, dispatches=[nf, nw]), None


def kernel_only(a, spec, X, kappa):
    """Run a few launches of one kernel (for hardware-counter passes) after a
    plain device copy of known size that calibrates the byte counters."""
    import torch
    import stfem_amd as st
    n_copy = COPY_BYTES // 8  # doubles
    src = torch.ones(n_copy, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    log(f"calibration copy: {COPY_BYTES} bytes read, {COPY_BYTES} bytes written per launch")
    g = st.Problem(dict(spec, mg_float=1), vertices=X, kappa=kappa)
    L = g.n_levels - 1
    for k in a.kernel_only.split(","):
        what = {"vanka": (6, L), "vmult32": (1, L), "vmult64": (0, L)}[k]
        ms = g.time_kernel(what[0], what[1], warmup=1, reps=5)
        byts = g.bytes_model(1, L) if k == "vanka" else g.bytes_model(0)
        log(f"{k}: {ms * 1e3:.1f} us per launch, model bytes {byts:.4e}")
    g.close()


def vmult_flops(dim, p, nt, n_cells):
    """Flop model of one space-time vmult: per cell and time component
    8 dim (p+1)^(dim+1) for the sum-factorised interpolation and gradients
    (in and out) plus (p+1)^dim (2 dim^2 + 2 + 2 nt (dim+1)) at the quadrature
    points (metric, coefficient, time mixing)."""
    n = p + 1
    per = 8 * dim * n ** (dim + 1) + n ** dim * (2 * dim * dim + 2 + 2 * nt * (dim + 1))
    return float(per) * nt * n_cells


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["headline", "q3", "wave"], default="headline",
                    help="headline: 64^3 cells, Q4 space, dG(3) time (67.9M dofs per slab); "
                         "q3: 32^3 cells, Q3/dG(2); wave: 3D wave, Q2/cG(2)")
    ap.add_argument("--refine", type=int, default=None,
                    help="cells per side 2^refine (headline 6, q3 and wave 5)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--pmc", choices=["on", "off"], default="on",
                    help="rocprofv3 counter passes behind roofline.traffic (two extra "
                         "processes that set up the FP32 multigrid problem and launch the "
                         "dominant kernel; about 50 s each for the 64^3 headline)")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="override a solver spec key (e.g. atomic_scatter=1)")
    ap.add_argument("--kernel-only", default=None, metavar="KERNELS",
                    help="comma-separated subset of vanka, vmult32, vmult64: set up the FP32 "
                         "multigrid problem and launch only these finest-level kernels")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(the multi-GPU decomposition is a design only)", file=sys.stderr)
        sys.exit(2)

    import torch
    torch.cuda.synchronize()
    if a.workload == "wave":
        return main_wave(a)
    make = workload_headline if a.workload == "headline" else workload_q3
    spec, X, kappa, desc = make(a.refine) if a.refine else make()
    for kv in a.set:
        k, v = kv.split("=", 1)
        spec[k] = int(v) if v.lstrip("-").isdigit() else float(v)
    if a.kernel_only:
        return kernel_only(a, spec, X, kappa)
    period = int(round(spec["end_time"] / (spec["nts"] * spec["tau"])))

    # FP32 multigrid inside FP64 FGMRES (headline)
    s32 = dict(spec, mg_float=1)
    g32, setup32, ms32, its32, ok32 = run_gpu(s32, X, kappa, a.steps, a.warmup)
    g32.spec_nts = spec["nts"]
    nst = g32.n_nodes * g32.nt
    dofs32 = a.steps * nst / (ms32 * 1e-3)

    # kernels: FP64 fine vmult, FP32 finest-level vmult and Vanka
    L = g32.n_levels - 1
    vm64_ms = g32.time_kernel(0, warmup=3, reps=20)
    vm64_bytes = g32.bytes_model(0)
    vm64_flops = vmult_flops(3, spec["p"], g32.nt, g32.n_cells)
    vm32_ms = g32.time_kernel(1, L, warmup=3, reps=20)
    nv = (g32.level_info(L)["nc0"] + 1) ** 3
    vm32_bytes = 4.0 * 2 * nst + 4.0 * g32.n_cells + 8.0 * 3 * nv
    # the sweep as the V-cycle runs it (sweep kernel + node gather), and the
    # sweep kernel alone (the roofline's kernel)
    vk_ms = g32.time_kernel(2, L, warmup=2, reps=10)
    vk_kernel_ms = g32.time_kernel(6, L, warmup=2, reps=10)
    vk_bytes = g32.bytes_model(1, L)
    # the transfer between the two finest levels
    tp_ms, tp_bytes = g32.time_kernel(4, L, warmup=2, reps=10), g32.bytes_model(2, L)
    tr_ms, tr_bytes = g32.time_kernel(5, L, warmup=2, reps=10), g32.bytes_model(3, L)
    u32 = None if a.no_fp64 else final_state(g32)
    n_levels = g32.n_levels
    g32.close()

    # plain FP64: one T = 1 run (at most `period` slabs) after one warm-up slab
    dofs64, its64, ms64, setup64, ok64, diff, steps64 = None, None, None, None, None, None, 0
    if not a.no_fp64:
        steps64 = min(a.steps, period)
        s64 = dict(spec, mg_float=0)
        g64, setup64, ms64, its64, ok64 = run_gpu(s64, X, kappa, steps64, min(a.warmup, 1))
        g64.spec_nts = spec["nts"]
        dofs64 = steps64 * nst / (ms64 * 1e-3)
        u64 = final_state(g64)
        g64.close()
        diff = ((u32 - u64).norm() / u64.norm()).item()
        log(f"FP32-MG vs FP64 solution after 2 slabs: relative L2 difference {diff:.2e}")

    # dominant kernel of an FGMRES iteration on the finest level: 2s Vanka
    # sweeps (s pre- and s post-smoothing steps), about as many FP32 level
    # vmults, one FP64 vmult
    s_steps = int(spec.get("smoothing_steps", 1))
    if 2 * s_steps * vk_ms >= max(2 * s_steps * vm32_ms, vm64_ms):
        dom = dict(kernel="vanka_factored_fp32", ms=vk_kernel_ms, bytes=vk_bytes, only="vanka",
                   symbol=r"k_vanka_factored_v")
    else:
        dom = dict(kernel="st_vmult_fp64", ms=vm64_ms, bytes=vm64_bytes, only="vmult64",
                   symbol=r"k_st_apply<3, \d, \d, double")
    achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9

    traffic, traffic_note = None, "counter passes off (--pmc off)"
    if a.pmc == "on":
        traffic, traffic_note = pmc_traffic(dom["only"], dom["symbol"], a.workload, a.refine,
                                            timeout_s=300 if a.workload == "headline" else 150)
        log(f"traffic: {traffic if traffic else traffic_note}")

    cpu = None if a.no_cpu else cpu_baseline(a.workload)
    out = {
        "metric": "space-time DoF/s per solve to 1e-8 (FP32-MG vs FP64) + vmult HBM GB/s",
        "value": dofs32,
        "unit": "space-time DoF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms32 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (FGMRES) + f32 (multigrid)",
        "data": "synthetic",
        "config": {"workload": desc, "space_time_dofs_per_slab": nst, "n_levels": n_levels,
                   "slabs_per_run": period, "vanka": "factored",
                   "secondary_workloads": "bench.py --workload q3 (32^3 Q3/dG(2)), "
                                          "--workload wave (32^3 Q2/cG(2) wave)"},
        "fp64_dofs_per_s": dofs64,
        "fp64_steps": steps64,
        "speedup_fp32mg_vs_fp64": (dofs32 / dofs64) if dofs64 else None,
        "fgmres_its_fp32mg": its32,
        "fgmres_its_fp64": its64,
        "converged": {"fp32mg": ok32, "fp64": ok64},
        "ms_per_fgmres_iteration": {"fp32mg": ms32 / its32 if its32 else None,
                                    "fp64": (ms64 / its64) if its64 else None},
        "setup_s": {"fp32mg": setup32, "fp64": setup64},
        # both solves stop at a 1e-8 relative FGMRES residual; at sizes the
        # oracle solves, the GPU slab solves match its FP64 solutions to 1e-12
        "fp32mg_vs_fp64_rel_l2_2slabs": diff,
        "fp32mg_vs_fp64_note": ("GPU FP32-multigrid solve vs GPU FP64 solve after two slabs; the "
                                "FP64 oracle cannot solve this size, its slab solves are compared "
                                "with the GPU's (1e-12) in the GPU tests at small sizes, this "
                                "discretisation at 4^3 cells"),
        "vmult_fp64": {"ms": vm64_ms, "GBps": vm64_bytes / (vm64_ms * 1e-3) / 1e9,
                       "hbm_frac": vm64_bytes / (vm64_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "model_flops": vm64_flops,
                       "TFLOPs": vm64_flops / (vm64_ms * 1e-3) / 1e12,
                       "fp64_frac": vm64_flops / (vm64_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                       "dofs_per_s": nst / (vm64_ms * 1e-3)},
        "vmult_fp32": {"ms": vm32_ms, "GBps": vm32_bytes / (vm32_ms * 1e-3) / 1e9},
        "vanka_fp32": {"ms": vk_ms, "GBps": vk_bytes / (vk_ms * 1e-3) / 1e9,
                       "kernel_ms": vk_kernel_ms,
                       "note": "ms: sweep kernel + node gather as in the V-cycle; kernel_ms: "
                               "the sweep kernel alone (roofline)"},
        "transfer_fp32": {"levels": f"{L - 1} <-> {L}",
                          "prolongate": {"ms": tp_ms, "GBps": tp_bytes / (tp_ms * 1e-3) / 1e9},
                          "restrict": {"ms": tr_ms, "GBps": tr_bytes / (tr_ms * 1e-3) / 1e9}},
        "roofline": {"kernel": dom["kernel"], "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None,
                     "algorithmic_bytes": dom["bytes"], "ms": dom["ms"],
                     "traffic_detail": traffic if traffic else traffic_note},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
