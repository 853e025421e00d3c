"""Benchmark: solved DOF/s (assemble + Jacobi-PCG to 1e-10 relative residual),
FP32-in-FP64 refinement vs plain FP64, plus the adjoint gradient time.

Workload (one "step"): the 128^3 C3D8 unit cube (2,097,152 elements,
6,440,067 DOF), E_e ~ lognormal(ln 1e9, 0.5) from a seeded counter-based
generator, nu = 0.3, z = 0 clamped by 1e12 penalties, pressure 1e6 on z = 1;
one step = assemble K into the fixed CSR pattern + PCG to ||r||/||b|| < 1e-10.
`value` is the mixed-precision DOF/s = n_dof / (t_assemble + t_solve); the
plain FP64 form is timed the same way beside it.  Inputs (mesh, pattern,
E field, loads, BCs) are resident on the GPU before the timed region; the
pattern / slot-map build is setup and reported separately.

Adjoint: the 96^3 SIMP cube (rho ~ U[0.3,1], E = 1e-3 + rho^3 1e9): forward
solve, adjoint solve and the element sensitivity kernel; "adjoint grad ms" is
adjoint solve + sensitivity.

roofline: the dominant kernel of the mixed step (the FP32 SpMV fused with
p.Ap; the symmetric block-diagonal view on this workload), algorithmic bytes
per launch / mean launch time from HIP events on the library's stream, with
the HBM traffic from the committed rocprofv3 counter summary beside it.
cpu_baseline: the FP64 CPU oracle (C, OpenMP, the box's CPU share of
threads) on the same 128^3 problem, not sampled: assembly five times (median
and min) and one full Jacobi-PCG from 0, timed to 1e-10 and continued to
1e-12; that solution is the reference for the GPU solutions' errors, and the
oracle's FP64 operator gives their true residuals.

A solve that does not converge fails the bench (exit status 3): no number is
published for it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 128]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ufc-fem-kernel_amd"))
sys.path.insert(0, ROOT)

METRIC = "solved DOF/s (assemble+PCG to 1e-10 res) mixed vs FP64; adjoint grad ms"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def spmv32_bytes(n_dof, nnz, n_nodes, index_bytes=4, n_penalty=0):
    """Algorithmic bytes of one FP32 SpMV.  Stencil-pattern view
    (index_bytes 1): the off-diagonal values (nnz - 9 n_nodes, the self blocks
    are not stored) + one pattern byte per node + the penalty diagonal of the
    n_penalty clamped DOFs; 16-bit offset / 4-byte id views: all nnz values +
    2 or 4 bytes per neighbour.  Both: node pointers + x + y.  The pattern
    table (at most 256 x 64 bytes, cache resident) is not counted.
    Symmetric block-diagonal view (index_bytes 0): each off-diagonal block
    once, (nnz - 9 n_nodes) / 2 values (its second use, transposed for the
    other node, is a cache re-read and not counted), the penalty diagonal,
    one flag bit per node, x + y; no node pointers."""
    if index_bytes == 0:
        return (nnz - 9 * n_nodes) // 2 * 4 + n_penalty * 4 + n_nodes // 8 + 2 * n_dof * 4
    if index_bytes == 1:
        head = (nnz - 9 * n_nodes) * 4 + n_nodes + n_penalty * 4
    else:
        head = nnz * 4 + (nnz // 9) * index_bytes
    return head + (n_nodes + 1) * 4 + 2 * n_dof * 4


SPMV_EVENT_EVERY = 64

# kernel name of the FP32 SpMV for each neighbour-index width
# (the symmetric view's kernel is k_spmv_sym<true, T> with T the layout's tile size)
SPMV32_KERNEL = {0: "k_spmv_sym<true, {tile}>", 1: "k_spmv<float, true, 2>", 2: "k_spmv<float, true, 1>",
                 4: "k_spmv<float, true, 0>"}
SPMV32_FORM = {0: "FP32 symmetric block-diagonal SpMV, x window in LDS, difference form",
               1: "FP32 stencil-pattern node-block SpMV, difference form",
               2: "FP32 node-block SpMV (16-bit offsets), difference form",
               4: "FP32 node-block SpMV (4-byte ids), difference form"}


def spmv64_bytes(n_dof, nnz):
    """Algorithmic bytes of one FP64 scalar-CSR SpMV: values + col_idx + row_ptr + x + y."""
    return nnz * (8 + 4) + (n_dof + 1) * 4 + 2 * n_dof * 8


def spmv64_block_bytes(n_dof, nnz, n_nodes):
    """Algorithmic bytes of one FP64 node-block SpMV: values + one 4-byte
    neighbour id per 3x3 block + node pointers + x + y."""
    return nnz * 8 + nnz // 9 * 4 + (n_nodes + 1) * 4 + 2 * n_dof * 8


def storage_vs_precision(ctx, pr, torch, x, ufc_hip, headline, fp64, check_step, replace_rtol):
    """The 4.2x of mixed over FP64 mixes two changes: precision (FP64 -> FP32
    values and vectors) and storage (a column index per value -> one id per
    block -> each off-diagonal block stored once).  This times the solve with
    one change at a time on the same values: FP64 on the node-block layout
    (the CSR's index bytes removed, same bits), then the mixed solve on the
    FP32 4-byte-id view (the same layout in FP32) and on the stencil-pattern
    view, beside the two headline forms."""
    import numpy as np
    from ufc_hip import _abi as A
    out = {"note": ("fp64_csr -> fp64_block: index bytes only (same bits); fp64_block -> "
                    "mixed_id32: precision only (same node-block layout); mixed_id32 -> "
                    "mixed_symdiag: each off-diagonal block stored once")}

    def timed(prec, label, steps=2):
        repl = replace_rtol if prec == ufc_hip.PREC_MIXED else 0.0
        run_step(ctx, pr, prec, 1e-10, torch, x, repl)
        rs = []
        for _ in range(steps):
            r, _ = run_step(ctx, pr, prec, 1e-10, torch, x, repl)
            check_step(label, r, "storage/precision step")
            rs.append(r)
        return {"ms": float(np.mean([r["wall_ms"] for r in rs])), "spmv_ms": rs[-1]["spmv_ms"],
                "iters": rs[-1]["niter"], "steps": steps}

    ctx.set_fp64_view(A.FP64_VIEW_BLOCK)
    b = timed(ufc_hip.PREC_FP64, "fp64_block")
    ctx.set_fp64_view(A.FP64_VIEW_CSR)
    b["spmv_bytes"] = spmv64_block_bytes(pr.n_dof, ctx.nnz, ctx.n_nodes)
    out["fp64_csr"] = {"ms": fp64["ms"], "spmv_ms": fp64["spmv_ms"], "iters": fp64["iters"],
                       "spmv_bytes": spmv64_bytes(pr.n_dof, ctx.nnz)}
    out["fp64_block"] = b
    log(f"fp64 node-block layout: {b['ms']:.1f} ms, spmv {b['spmv_ms']:.4f} ms")
    nbc = int(np.unique(pr.bc_eq).size)
    for view, name, ib in ((A.FP32_VIEW_ID32, "mixed_id32", 4),
                           (A.FP32_VIEW_PATTERN, "mixed_pattern", 1)):
        ctx.set_fp32_view(view)
        r = timed(ufc_hip.PREC_MIXED, name)
        r["spmv_bytes"] = spmv32_bytes(pr.n_dof, ctx.nnz, ctx.n_nodes, ib, nbc)
        out[name] = r
        log(f"{name}: {r['ms']:.1f} ms, spmv {r['spmv_ms']:.4f} ms")
    ctx.set_fp32_view(-1)
    out["mixed_symdiag"] = {"ms": headline["ms"], "spmv_ms": None,
                            "iters": headline["inner_iters"]}
    out["fp64_csr_over_fp64_block"] = fp64["ms"] / b["ms"]
    out["fp64_block_over_mixed_id32"] = b["ms"] / out["mixed_id32"]["ms"]
    out["mixed_id32_over_mixed_symdiag"] = out["mixed_id32"]["ms"] / headline["ms"]
    return out


def measured_traffic(n_dof, kernel):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 counter summary, when it was taken on this workload."""
    path = os.path.join(ROOT, "profiles", "spmv_pcg128_hbm_traffic.json")
    if n_dof != 6440067 or not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f)["kernels"]
    return k[kernel]["traffic_bytes"] if kernel in k else None


def measured_assembly_flop(n_elems):
    """Executed FLOPs per assembly launch from the committed rocprofv3 VALU
    counter summary, when it was taken on this workload."""
    path = os.path.join(ROOT, "profiles", "assembly128_valu_flops.json")
    if n_elems != 2097152 or not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f)["kernels"]
    names = ("k_assemble<float, 0, 1>", "k_assemble<double, 0, 1>")
    if not all(n in k for n in names):  # summary of another build's kernels
        return None
    return {"fp32": k[names[0]]["executed_flop"], "fp64": k[names[1]]["executed_flop"]}


def run_step(ctx, pr, precision, rtol, torch, x, replace_rtol=0.0):
    import ufc_hip
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.assemble_k(pr.young_dev, pr.nu, precision=precision)
    t1 = time.perf_counter()
    u, info = ctx.pcg(precision=precision, rtol=rtol, inner_rtol=1e-5, x=x,
                      replace_rtol=replace_rtol)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    t = ctx.timing()
    return dict(wall_ms=(t2 - t0) * 1e3, assemble_wall_ms=(t1 - t0) * 1e3,
                assemble_ms=t.assemble_ms, solve_ms=t.solve_ms, spmv_ms=t.spmv_ms_avg,
                spmv_calls=t.spmv_calls, niter=info.niter, n_outer=info.n_outer,
                n_replace=t.n_replace,
                rnorm=info.rnorm, converged=info.converged,
                status=info.status), u


def cpu_threads():
    """Threads for the CPU oracle: OMP_NUM_THREADS when set (the box's CPU
    share), else every core of the affinity mask; never more than the mask."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else aff
    return max(1, min(n, aff)), aff


def cpu_baseline(pr, gpu_sols, repeats=5, rtol_full=1e-12):
    """FP64 oracle (C, OpenMP) on the same problem: pattern (setup), assembly
    `repeats` times (median and min), then one Jacobi-PCG from x = 0 to
    rtol_full that also records the time at which it crossed 1e-10 -- the
    measured CPU solve time of the benchmark's tolerance -- and whose
    solution is the reference for the GPU solutions' errors."""
    import numpy as np
    from oracle import oracle as O
    threads, aff = cpu_threads()
    t0 = time.perf_counter()
    rp, ci = O.csr_pattern(pr.coords.shape[0], pr.conn)
    t_pattern = time.perf_counter() - t0
    t_asm = []
    for k in range(repeats):
        t0 = time.perf_counter()
        val = O.assemble(pr.conn, pr.coords, pr.young, pr.nu, rp, ci, nthreads=threads)
        t_asm.append(time.perf_counter() - t0)
        log(f"cpu assembly {k}: {t_asm[-1]:.2f} s")
    F = np.zeros(pr.n_dof)
    O.face_loads_pressure(pr.load_elem, pr.load_face, pr.conn, pr.coords, pr.load_value[0], F)
    O.apply_penalty(pr.bc_eq, pr.bc_val, 1e12, rp, ci, val, F)
    log(f"cpu PCG to {rtol_full:g} ({threads} threads)")
    t0 = time.perf_counter()
    u_cpu, st, itf, rnf, it_mark, t_mark = O.pcg_mark(rp, ci, val, F, rtol_full, 1e-10,
                                                      nthreads=threads, report_every=250)
    t_full = time.perf_counter() - t0
    if st != 0 or it_mark < 0:
        raise RuntimeError(f"CPU oracle PCG did not converge (status {st}, {itf} iterations)")
    t0 = time.perf_counter()
    _, _, it1, _ = O.pcg(rp, ci, val, F, rtol=1e-10, max_iter=3, nthreads=1)
    per_it1 = (time.perf_counter() - t0) / max(it1, 1)
    asm_med = float(np.median(t_asm))
    out = {"value": pr.n_dof / (asm_med + t_mark), "unit": "DOF/s", "cores": threads,
           "kind": "port",
           "sample": (f"FP64 C oracle (OpenMP, {threads} threads; nproc {os.cpu_count()}, "
                      f"affinity {aff}) on the same {pr.nx}^3 problem, not sampled: assembly "
                      f"x{repeats} (median {asm_med:.2f} s, min {min(t_asm):.2f} s), one "
                      f"Jacobi-PCG from 0 to 1e-10 ({it_mark} iterations, {t_mark:.1f} s, "
                      f"measured inside a run continued to {rtol_full:g}); CSR pattern "
                      f"{t_pattern:.2f} s is setup and not counted"),
           "threads": threads, "nproc": os.cpu_count(), "affinity_cores": aff,
           "assemble_s_median": asm_med, "assemble_s_min": float(min(t_asm)),
           "assemble_repeats": repeats, "solve_s": t_mark, "solve_iters": int(it_mark),
           "pcg_iter_ms": t_mark / it_mark * 1e3, "pcg_iter_ms_1core": per_it1 * 1e3,
           "pattern_s": t_pattern,
           "full_solve": {"rtol": rtol_full, "iters": int(itf), "s": t_full}}
    errs = {}
    bn = np.linalg.norm(F)
    for name, u in gpu_sols.items():
        errs[f"{name}_vs_cpu_oracle_rel"] = float(np.linalg.norm(u - u_cpu) /
                                                  np.linalg.norm(u_cpu))
        r = F - O.spmv(rp, ci, val, u, nthreads=threads)
        errs[f"{name}_oracle_true_residual"] = float(np.linalg.norm(r) / bn)
    errs["cpu_oracle_rtol"] = rtol_full
    return out, errs


def cpu_second_opinion(repeats=3):
    """The small reference configuration (20x4x4 cantilever, 1575 DOF) timed
    two ways on the host: oracle/_ref/ref_cantilever, a Fortran program built
    from the reference's own compiled sparse module (its COO element assembly
    and CSR conversion, and NM_CSR_MatVec_Optimized inside the PCG), and the C
    oracle (1 thread); both assemble and solve to 1e-10, `repeats` times
    (median and min), and their displacements are compared.  None when the
    reference-built program is absent."""
    import subprocess
    import tempfile
    import numpy as np
    from oracle import oracle as O
    from ufc_hip import problems as P
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_cantilever")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "u.txt")
        subprocess.run([exe, str(repeats), out], check=True, timeout=600)
        txt = open(out).read().split()
    ndof, nnz, iters, rep = (int(v) for v in txt[:4])
    vals = np.array([float(v) for v in txt[4:]])
    t_asm, t_sol, u_f = vals[:rep], vals[rep:2 * rep], vals[2 * rep:]
    pr = P.cantilever()
    t_oa, t_os = [], []
    for _ in range(max(repeats, 5)):
        t0 = time.perf_counter()
        rp, ci = O.csr_pattern(pr.coords.shape[0], pr.conn)
        val = O.assemble(pr.conn, pr.coords, pr.young, pr.nu, rp, ci, nthreads=1)
        F = np.zeros(pr.n_dof)
        O.face_loads_traction(pr.load_elem, pr.load_face, pr.conn, pr.coords, pr.load_value, F)
        O.apply_penalty(pr.bc_eq, pr.bc_val, 1e12, rp, ci, val, F)
        t1 = time.perf_counter()
        uo, st, it, _ = O.pcg(rp, ci, val, F, rtol=1e-10, nthreads=1)
        t_os.append(time.perf_counter() - t1)
        t_oa.append(t1 - t0)
    return {"workload": "20x4x4 C3D8 cantilever (1575 DOF), assemble + Jacobi-PCG to 1e-10",
            "fortran_reference_kernels": {
                "assemble_s_median": float(np.median(t_asm)), "assemble_s_min": float(t_asm.min()),
                "solve_s_median": float(np.median(t_sol)), "solve_s_min": float(t_sol.min()),
                "iters": iters, "nnz": nnz, "repeats": rep, "cores": 1,
                "note": "reference NM_COO_AddElementMatrix (a linear search of the COO list per "
                        "entry) + NM_COO_to_CSR, NM_CSR_MatVec_Optimized; element matrices, "
                        "loads, penalty and the PCG loop restated in Fortran"},
            "c_oracle": {"assemble_s_median": float(np.median(t_oa)),
                         "solve_s_median": float(np.median(t_os)), "iters": int(it),
                         "cores": 1, "repeats": len(t_oa)},
            "u_rel_diff": float(np.linalg.norm(u_f - uo) / np.linalg.norm(uo))}


# J2 update per Gauss point, algorithmic: committed state read and trial
# state written (13 doubles each), the factored tangent written (mu', lam', c,
# n(6): 9 doubles), flag (4 B); per element its connectivity, per node
# coordinates, u and u_n
J2_BYTES_PER_GP = 13 * 8 * 2 + 9 * 8 + 4
# FLOPs per Gauss point: J = dN/dxi x (9 x 8 FMA = 144), det + inverse (50),
# dN/dx = J^-T dN/dxi (8 x 3 x 5 = 120), strain increment from 8 nodes
# (8 x 27 = 216), elastic predictor, deviator and q (40), radial return and
# state update (40), factored consistent tangent (10)
J2_FLOP_PER_GP = 144 + 50 + 120 + 216 + 40 + 40 + 10


def j2_roofline(n_elems, n_nodes, ms):
    """HBM roofline of one k_j2_update launch (the bound: ~0.6 FLOP per byte
    against FP64 VALU's ~10 per byte of HBM)."""
    ngp = 8 * n_elems
    b = ngp * J2_BYTES_PER_GP + n_elems * 32 + n_nodes * 72
    f = ngp * J2_FLOP_PER_GP
    ach = b / (ms * 1e6)
    return {"kernel": "k_j2_update", "bound": "hbm", "bytes_per_launch": b, "ms_per_launch": ms,
            "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "flop_per_launch": f, "fp64_tflops": f / (ms * 1e9),
            "fp64_frac_of_78.6": f / (ms * 1e9) / 78.6}


def cpu_adjoint_baseline(pa, young_gpu=None, u_gpu=None, grads=None):
    """Config-4 workload on the FP64 oracle (threads as cpu_baseline), not
    sampled: assembly, forward Jacobi-PCG to 1e-10, the adjoint solve
    K^T lambda = f (NM_Adjoint_Solve on symmetric K, Jacobi-PCG from 0 to
    1e-10) and the element sensitivities.  With the GPU's E field, forward
    solution and gradients: the oracle's FP64 residual of the GPU's u, the
    oracle's sensitivities from the GPU's u (lambda = u for compliance)
    against the GPU gradient, and each GPU gradient against the oracle's
    (errors, outside the timed legs)."""
    import numpy as np
    from oracle import oracle as O
    threads, aff = cpu_threads()
    rp, ci = O.csr_pattern(pa.coords.shape[0], pa.conn)
    t0 = time.perf_counter()
    young = pa.young if young_gpu is None else young_gpu
    val = O.assemble(pa.conn, pa.coords, young, pa.nu, rp, ci, nthreads=threads)
    t1 = time.perf_counter()
    F = np.zeros(pa.n_dof)
    O.face_loads_pressure(pa.load_elem, pa.load_face, pa.conn, pa.coords, pa.load_value[0], F)
    f0 = F.copy()
    O.apply_penalty(pa.bc_eq, pa.bc_val, 1e12, rp, ci, val, F)
    log("cpu adjoint workload: forward PCG")
    u, st, it_f, _, _, _ = O.pcg_mark(rp, ci, val, F, 1e-10, 0.0, nthreads=threads,
                                      report_every=500)
    t2 = time.perf_counter()
    log("cpu adjoint workload: adjoint PCG")
    lam, st2, it_a, _, _, _ = O.pcg_mark(rp, ci, val, f0, 1e-10, 0.0, nthreads=threads,
                                         report_every=500)
    t3 = time.perf_counter()
    rho = pa.extra["rho"]
    g = O.sensitivity(pa.conn, pa.coords, pa.nu, 3.0 * rho ** 2 * 1e9, u, lam, nthreads=threads)
    t4 = time.perf_counter()
    if st != 0 or st2 != 0:
        raise RuntimeError("CPU oracle adjoint workload did not converge")
    err = {}
    if u_gpu is not None:
        err["gpu_u_oracle_residual"] = float(np.linalg.norm(F - O.spmv(rp, ci, val, u_gpu,
                                                                       threads)) /
                                             np.linalg.norm(F))
        err["gpu_u_vs_cpu_oracle_rel"] = float(np.linalg.norm(u_gpu - u) / np.linalg.norm(u))
    for name, gg in (grads or {}).items():
        if gg is None:
            continue
        err[f"grad_{name}_vs_cpu_oracle_rel"] = float(np.abs(gg - g).max() / np.abs(g).max())
        if u_gpu is not None and name == "mixed":
            go = O.sensitivity(pa.conn, pa.coords, pa.nu, 3.0 * rho ** 2 * 1e9, u_gpu, u_gpu,
                               nthreads=threads)
            err["grad_mixed_vs_oracle_sensitivity_of_gpu_u"] = float(np.abs(gg - go).max() /
                                                                     np.abs(go).max())
    return {"kind": "port", "cores": threads,
            "sample": (f"FP64 C oracle ({threads} threads) on the same {pa.nx}^3 SIMP cube, not "
                       "sampled: assembly, forward PCG to 1e-10, adjoint PCG to 1e-10, "
                       "sensitivity kernel"),
            "forward_ms": (t2 - t0) * 1e3, "grad_ms": (t4 - t2) * 1e3,
            "assemble_ms": (t1 - t0) * 1e3, "forward_iters": int(it_f),
            "adjoint_iters": int(it_a), "sensitivity_ms": (t4 - t3) * 1e3,
            "grad_checksum": float(np.abs(g).sum()), "errors": err}


def cpu_j2_baseline(pj, props, u_end, n_newton, n_linear):
    """Config-5 workload on the FP64 oracle, a bounded sample projected to
    the GPU job's counts: one J2 update + internal force over all Gauss
    points (a strain increment from the unloaded state to the job's final
    displacement, every point on the plastic branch), one tangent assembly
    and 20 Jacobi-PCG iterations on that tangent, all with the CPU share of
    threads; CPU job = n_newton x (update + assembly) + n_linear x PCG
    iteration."""
    import numpy as np
    from oracle import oracle as O
    threads, aff = cpu_threads()
    nn = pj.coords.shape[0]
    rp, ci = O.csr_pattern(nn, pj.conn)
    s0 = np.zeros(pj.n_elems * 8 * 13)
    t0 = time.perf_counter()
    nf, st, Dgp, Fi = O.j2_internal_force(nn, pj.conn, pj.coords, props, u_end,
                                          np.zeros(pj.n_dof), s0, nthreads=threads)
    t1 = time.perf_counter()
    val = O.assemble_dgp(pj.conn, pj.coords, Dgp, rp, ci, nthreads=threads)
    t2 = time.perf_counter()
    O.apply_penalty(pj.bc_eq, pj.bc_val, 1e12, rp, ci, val, None)
    t3 = time.perf_counter()
    _, _, it, _ = O.pcg(rp, ci, val, Fi, rtol=1e-30, max_iter=20, nthreads=threads)
    t4 = time.perf_counter()
    per_it = (t4 - t3) / max(it, 1)
    job = n_newton * ((t1 - t0) + (t2 - t1)) + n_linear * per_it
    return {"kind": "port", "cores": threads,
            "sample": (f"FP64 C oracle ({threads} threads) on the same {pj.nx}^3 mesh: one J2 "
                       f"update + internal force ({(t1 - t0) * 1e3:.0f} ms, "
                       f"{int(np.count_nonzero(st.reshape(-1, 13)[:, 6]))} plastic points), one "
                       f"tangent assembly ({(t2 - t1) * 1e3:.0f} ms), {it} PCG iterations "
                       f"({per_it * 1e3:.2f} ms each); projected to the GPU job's "
                       f"{n_newton} Newton and {n_linear} linear iterations"),
            "update_ms": (t1 - t0) * 1e3, "assemble_ms": (t2 - t1) * 1e3,
            "pcg_iter_ms": per_it * 1e3, "job_s_projected": job}


def oracle_j2_replay(pj, props, target, u_incs, state_gpu):
    """The oracle's replay of a GPU J2 job: from the GPU's converged u of
    every increment, orc_j2_internal_force recomputes the J2 state from the
    previous increment's and the internal force; the residual R = P ubar -
    F_int - P u of each increment relative to ||F_int|| (the GPU's Newton
    stops at 1e-8 by its own FP64 residual), and the GPU's final state
    (stresses; equivalent and plastic strains) against the replayed one."""
    import numpy as np
    from oracle import oracle as O
    from ufc_hip.nonlinear import to_host
    threads, _ = cpu_threads()
    nn = pj.coords.shape[0]
    state = np.zeros(pj.n_elems * 8 * 13)
    u_prev = np.zeros(pj.n_dof)
    rels, fails = [], 0
    n = len(u_incs)
    for k, uk in enumerate(u_incs, start=1):
        uk = to_host(uk)
        nf, state, _, fi = O.j2_internal_force(nn, pj.conn, pj.coords, props, uk, u_prev, state,
                                               nthreads=threads)
        fails += int(nf)
        r = -fi
        r[pj.bc_eq] += 1e12 * (target * (k / n) - uk[pj.bc_eq])
        rels.append(float(np.linalg.norm(r) / np.linalg.norm(fi)))
        u_prev = uk
    so, sg = state.reshape(-1, 13), state_gpu.reshape(-1, 13)
    return {"oracle_residual_rel_max": max(rels), "oracle_residual_rel": rels,
            "local_failures": fails,
            "stress_vs_oracle_replay_rel": float(np.abs(sg[:, :6] - so[:, :6]).max() /
                                                 np.abs(so[:, :6]).max()),
            "plastic_strain_vs_oracle_replay_rel": float(np.abs(sg[:, 6:] - so[:, 6:]).max() /
                                                         np.abs(so[:, 6:]).max())}


def renumbered_run(args, torch, dev, ufc_hip, P, n, sols):
    """The same cube with its node ids permuted at random (seeded): the
    library recovers the box's lexicographic numbering for the symmetric
    FP32 view; mixed and FP64 steps are timed as above (2 and 1 steps after
    one warm-up each) and the solutions compared with the natural
    numbering's after un-permuting."""
    import numpy as np
    pr = P.lognormal_cube(n)
    perm = np.random.default_rng(P.SEED + 5).permutation(pr.coords.shape[0])
    coords = np.empty_like(pr.coords)
    coords[perm] = pr.coords
    pr.coords = coords
    pr.conn = perm[pr.conn].astype(np.int32)
    pr.bc_eq = (3 * perm[pr.bc_eq // 3] + pr.bc_eq % 3).astype(np.int32)
    log(f"renumbered {n}^3: setup")
    t0 = time.perf_counter()
    cr = ufc_hip.Context(0)
    P.setup_context(cr, pr, ufc_hip.PREC_MIXED)
    cr.enable_timing(SPMV_EVENT_EVERY)
    t1 = time.perf_counter()
    pr.young_dev = torch.as_tensor(pr.young, device=dev)
    x = torch.zeros(pr.n_dof, dtype=torch.float64, device=dev)
    res = {"workload": (f"the {n}^3 cube with node ids permuted at random (seed "
                        f"{P.SEED + 5}); the library finds the box numbering"),
           "fp32_view": int(cr.fp32_view), "nb_index_bytes": int(cr.nb_index_bytes),
           "setup_s": t1 - t0}
    for name, prec, repl, k in (("mixed", ufc_hip.PREC_MIXED, args.replace_rtol, 2),
                                ("fp64", ufc_hip.PREC_FP64, 0.0, 1)):
        run_step(cr, pr, prec, 1e-10, torch, x, repl)
        runs = []
        for _ in range(k):
            r, u = run_step(cr, pr, prec, 1e-10, torch, x, repl)
            if not r["converged"] or not r["rnorm"] < 1e-10:
                raise RuntimeError(f"renumbered {name} solve did not converge")
            runs.append(r)
        ms = float(np.mean([r["wall_ms"] for r in runs]))
        u_nat = u.detach().cpu().numpy().reshape(-1, 3)[perm].ravel()
        ref = sols.get(name)
        res[name] = {"ms": ms, "iters": runs[-1]["niter"], "spmv_ms": runs[-1]["spmv_ms"]}
        if ref is not None:
            res[name]["vs_natural_numbering_rel"] = float(np.linalg.norm(u_nat - ref) /
                                                          np.linalg.norm(ref))
        log(f"renumbered {name}: {ms:.1f} ms, it {runs[-1]['niter']}")
    res["mixed_over_fp64"] = res["fp64"]["ms"] / res["mixed"]["ms"]
    cr.close()
    return res


def shard_run(args, torch, dev, ufc_hip, P, n, check_step):
    """One GPU's share of the n^3 cube split into 8 z-slabs (n x n x n/8
    hexes, the same seeded E field; z=0 clamped, pressure on the slab's top):
    what a GPU of an 8-GPU run computes per step, without the halo exchange
    and the two all-reduces per iteration that only a multi-GPU run has.
    Mixed and FP64 steps timed as the headline (3 steps after one warm-up),
    the FP32 SpMV's launch time (events around every 8th launch) and rate."""
    import numpy as np
    pr = P.slab_share(n, 8)
    log(f"shard {pr.name}: setup")
    ctx = ufc_hip.Context(0)
    P.setup_context(ctx, pr, ufc_hip.PREC_MIXED)
    ctx.enable_timing(8)
    pr.young_dev = torch.as_tensor(pr.young, device=dev)
    x = torch.zeros(pr.n_dof, dtype=torch.float64, device=dev)
    res = {"workload": (f"{pr.name}: the bottom of 8 z-slabs of the {n}^3 lognormal cube "
                        "(same E field), z=0 clamped, p=1e6 on the slab top; no halo exchange "
                        "or all-reduce (one GPU)"),
           "n_dof": pr.n_dof, "nnz": ctx.nnz}
    for name, prec, repl in (("mixed", ufc_hip.PREC_MIXED, args.replace_rtol),
                             ("fp64", ufc_hip.PREC_FP64, 0.0)):
        r, _ = run_step(ctx, pr, prec, 1e-10, torch, x, repl)
        check_step(f"shard {name}", r, "warmup")
        runs = []
        for k in range(3):
            r, _ = run_step(ctx, pr, prec, 1e-10, torch, x, repl)
            check_step(f"shard {name}", r, f"step {k}")
            runs.append(r)
        ms = float(np.mean([q["wall_ms"] for q in runs]))
        it = runs[-1]["niter"]
        res[name] = {"ms": ms, "dof_per_s": pr.n_dof / (ms / 1e3), "iters": it,
                     "us_per_iteration": runs[-1]["solve_ms"] / max(it, 1) * 1e3,
                     "spmv_us": runs[-1]["spmv_ms"] * 1e3}
        log(f"shard {name}: {ms:.1f} ms, it {it}, "
            f"{res[name]['us_per_iteration']:.1f} us/iteration, spmv "
            f"{res[name]['spmv_us']:.1f} us")
    b32 = spmv32_bytes(pr.n_dof, ctx.nnz, ctx.n_nodes, ctx.nb_index_bytes,
                       int(np.unique(pr.bc_eq).size))
    res["mixed"]["spmv_gbs"] = b32 / (res["mixed"]["spmv_us"] * 1e3)
    res["mixed"]["spmv_frac"] = res["mixed"]["spmv_gbs"] / HBM_PEAK_GBS
    res["fp64"]["spmv_gbs"] = spmv64_bytes(pr.n_dof, ctx.nnz) / (res["fp64"]["spmv_us"] * 1e3)
    res["mixed_over_fp64"] = res["fp64"]["ms"] / res["mixed"]["ms"]
    ctx.close()
    return res


def copy_bandwidth(torch, dev, nbytes=1 << 31):
    """Device-to-device copy of nbytes (reads and writes nbytes): the
    attainable-bandwidth denominator beside the 8 TB/s peak."""
    a = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(5):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    del a, b
    return 2.0 * nbytes / (ms * 1e6)


def run_j2(args, out, P, ufc_hip, torch):
    """Config 5: the 64^3 jittered J2 job, inexact and exact Newton, into out["j2"]."""
    import numpy as np
    from ufc_hip.nonlinear import j2_static
    nj = args.j2_size
    log(f"J2 {nj}^3 jittered, 10 increments to 0.5% strain")
    pj = P.j2_cube(nj)
    target = np.zeros(pj.bc_eq.size)
    target[pj.extra["n_clamp"]:] = pj.extra["strain"]
    with ufc_hip.Context(0) as cj:
        cj.build_pattern(pj.coords, pj.conn)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = j2_static(cj, pj.bc_eq, target, pj.extra["props"], n_increments=10,
                        precision=ufc_hip.PREC_MIXED, keep_increments=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        state_inexact = cj.j2_get_state() if not args.no_cpu else None
        # the same inexact-Newton job with the extrapolated first iterate
        cj.build_pattern(pj.coords, pj.conn)
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        res_pr = j2_static(cj, pj.bc_eq, target, pj.extra["props"], n_increments=10,
                           precision=ufc_hip.PREC_MIXED, keep_increments=True, predictor=True)
        torch.cuda.synchronize()
        t5 = time.perf_counter()
        state_pred = cj.j2_get_state() if not args.no_cpu else None
        # exact Newton (linear solves to 1e-10) on the same job: the
        # consistent tangent's iteration count per increment
        cj.build_pattern(pj.coords, pj.conn)
        t2 = time.perf_counter()
        res_ex = j2_static(cj, pj.bc_eq, target, pj.extra["props"], n_increments=10,
                           precision=ufc_hip.PREC_MIXED, lin_rtol=1e-10, keep_increments=True)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        state_exact = cj.j2_get_state() if not args.no_cpu else None
    if not (res.converged and res_ex.converged and res_pr.converged):
        print(json.dumps({"metric": METRIC, "error": "the J2 Newton job did not converge"}),
              flush=True)
        sys.exit(3)
    out["j2"] = {"workload": (f"{nj}^3 C3D8 cube, interior nodes jittered +-10% h, "
                              "E 200e9, nu 0.3, sigma_y 250e6, H 2e9, u_z on z=1 to "
                              "0.5% strain in 10 increments; inexact Newton on the "
                              "consistent tangent to |R| <= 1e-8 |F_int| (linear tolerance "
                              "1e-6 on the first iteration, then Eisenstat-Walker <= 1e-3), "
                              "inner solve FP32-in-FP64"),
                 "n_dof": pj.n_dof, "converged": res.converged, "total_ms": (t1 - t0) * 1e3,
                 "newton_iters": [i.newton_iters for i in res.increments],
                 "linear_iters": int(sum(i.linear_iters for i in res.increments)),
                 "yielded_gauss_points_final": int(res.increments[-1].n_yield)
                 if res.increments else 0,
                 "exact_newton": {"linear_rtol": 1e-10, "total_ms": (t3 - t2) * 1e3,
                                  "newton_iters": [i.newton_iters for i in res_ex.increments],
                                  "rnorm_rel": [i.rnorm_rel for i in res_ex.increments],
                                  "history_last_increment":
                                      res_ex.increments[-1].history,
                                  "linear_iters": int(sum(i.linear_iters
                                                          for i in res_ex.increments))},
                 "u_exact_vs_inexact_rel": float(np.linalg.norm(res.u - res_ex.u) /
                                                 np.linalg.norm(res_ex.u)),
                 "predictor": {"form": ("the timed job with the first Newton iterate of each "
                                        "increment extrapolated from the previous increment "
                                        "(u_n + dlam/dlam_prev (u_n - u_n-1)) instead of u_n; "
                                        "its first linear solve at eta_max instead of 1e-6"),
                               "total_ms": (t5 - t4) * 1e3,
                               "newton_iters": [i.newton_iters for i in res_pr.increments],
                               "linear_iters": int(sum(i.linear_iters
                                                       for i in res_pr.increments)),
                               "u_vs_inexact_rel": float(np.linalg.norm(res_pr.u - res.u) /
                                                         np.linalg.norm(res.u))}}
    log(f"J2 predictor: {(t5 - t4):.2f} s, newton {out['j2']['predictor']['newton_iters']}, "
        f"linear {out['j2']['predictor']['linear_iters']}")
    log(f"J2: {(t1 - t0):.2f} s, newton {out['j2']['newton_iters']}")
    upd = res.j2_update_ms + res_ex.j2_update_ms
    out["j2"]["roofline"] = j2_roofline(pj.n_elems, pj.coords.shape[0],
                                        float(np.median(upd)))
    out["j2"]["roofline"]["launches"] = len(upd)
    out["j2"]["tangent_assembly_ms_median"] = float(np.median(res.assemble_ms))
    log(f"J2 update {np.median(upd):.3f} ms, tangent assembly "
        f"{np.median(res.assemble_ms):.3f} ms")
    if not args.no_cpu:
        log("J2: oracle replay of both jobs")
        out["j2"]["errors"] = {
            "inexact": oracle_j2_replay(pj, pj.extra["props"], target, res.u_increments,
                                        state_inexact),
            "exact": oracle_j2_replay(pj, pj.extra["props"], target, res_ex.u_increments,
                                      state_exact),
            "predictor": oracle_j2_replay(pj, pj.extra["props"], target, res_pr.u_increments,
                                          state_pred)}
        log(f"J2 oracle replay: {json.dumps(out['j2']['errors'])[:400]}")
        n_newton = int(sum(i.newton_iters for i in res.increments))
        cb = cpu_j2_baseline(pj, pj.extra["props"], res.u, n_newton,
                             out["j2"]["linear_iters"])
        cb["value_job_s_gpu"] = (t1 - t0)
        out["j2"]["cpu_baseline"] = cb
        log(f"cpu J2: {cb}")



def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--adjoint-size", type=int, default=96)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU oracle leg (baseline timing and the full-size errors)")
    ap.add_argument("--compare-steps", type=int, default=5,
                    help="timed steps of the comparison forms (FP64, restarted mixed)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--replace-rtol", type=float, default=0.1,
                    help="mixed: FP64 residual replacement inside the inner solve every this "
                         "reduction, search direction kept (>= 1: restarted refinement steps)")
    ap.add_argument("--no-restarted", action="store_true",
                    help="skip the comparison run of the restarted refinement form")
    ap.add_argument("--fd-check", type=int, default=8,
                    help="central-difference check of the adjoint gradient on N elements "
                         "(0: none)")
    ap.add_argument("--no-j2", action="store_true", help="skip the J2 Newton workload")
    ap.add_argument("--no-views", action="store_true",
                    help="skip the storage-vs-precision runs (FP64 node-block layout, FP32 "
                         "4-byte-id and stencil-pattern views)")
    ap.add_argument("--no-renumbered", action="store_true",
                    help="skip the randomly renumbered copy of the cube")
    ap.add_argument("--no-shard", action="store_true",
                    help="skip one GPU's share of an 8-way z-slab split of the cube")
    ap.add_argument("--j2-size", type=int, default=64)
    ap.add_argument("--j2-only", action="store_true",
                    help="run only the J2 Newton workload (profiling); prints its JSON")
    ap.add_argument("--pmc-sample", action="store_true",
                    help="short run for hardware-counter passes: 10 PCG iterations in each "
                         "precision plus a 2 GiB device copy for calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the library drives one "
              "GPU per context and a multi-GPU split of the mesh is not implemented or measured.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import ufc_hip
    from ufc_hip import problems as P

    dev = torch.device("cuda:0")
    if args.j2_only:
        out = {"metric": METRIC, "note": "J2 workload only (--j2-only); not the headline"}
        run_j2(args, out, P, ufc_hip, torch)
        print(json.dumps(out), flush=True)
        return
    n = args.size
    log(f"setup {n}^3")
    t0 = time.perf_counter()
    pr = P.lognormal_cube(n)
    t1 = time.perf_counter()
    ctx = ufc_hip.Context(0)
    P.setup_context(ctx, pr, ufc_hip.PREC_MIXED)
    pr.young_dev = torch.as_tensor(pr.young, device=dev)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    log(f"mesh {t1 - t0:.2f} s, pattern+slot map+loads+BC {t2 - t1:.2f} s, nnz {ctx.nnz}")
    x = torch.zeros(pr.n_dof, dtype=torch.float64, device=dev)
    # HIP event pairs around every 64th SpMV launch of the timed steps (an
    # event record costs ~5.6 us of queue time; sampling keeps the timed
    # steps within ~0.05% of untimed ones; ~40 samples per mixed step)
    ctx.enable_timing(SPMV_EVENT_EVERY)
    if args.pmc_sample:
        # calibration: a plain copy of 2 GiB (reads 2 GiB, writes 2 GiB)
        a = torch.ones(2 ** 28, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        for prec in (ufc_hip.PREC_MIXED, ufc_hip.PREC_FP64):
            ctx.assemble_k(pr.young_dev, pr.nu, precision=prec)
            ctx.pcg(precision=prec, rtol=1e-10, max_iter=10, max_outer=1, x=x)
        torch.cuda.synchronize()
        log("pmc sample done")
        return

    res = {}
    sols = {}
    modes = [("mixed", ufc_hip.PREC_MIXED, args.replace_rtol)]
    if not args.no_fp64:
        modes.append(("fp64", ufc_hip.PREC_FP64, 0.0))
    if not args.no_restarted:
        modes.append(("mixed_restarted", ufc_hip.PREC_MIXED, 1.0))
    failed = []

    def check_step(name, r, phase):
        # a step that did not reach the tolerance is not a solve: the bench fails
        if not r["converged"] or r["status"] != 0 or not r["rnorm"] < 1e-10:
            failed.append({"mode": name, "phase": phase, "status": r["status"],
                           "converged": bool(r["converged"]), "rnorm": r["rnorm"]})
            log(f"NOT CONVERGED: {name} {phase}: status {r['status']}, rnorm {r['rnorm']:.3e}")

    for name, prec, repl in modes:
        # the headline (mixed) runs --steps/--warmup; the comparison forms
        # run at most --compare-steps after one warm-up step
        n_steps = args.steps if name == "mixed" else min(args.steps, args.compare_steps)
        n_warm = args.warmup if name == "mixed" else min(args.warmup, 1)
        for w in range(n_warm):
            r, _ = run_step(ctx, pr, prec, 1e-10, torch, x, repl)
            check_step(name, r, f"warmup {w}")
            log(f"warmup {name}: {r['wall_ms']:.1f} ms, it {r['niter']}, outer {r['n_outer']}")
        runs = []
        for k in range(n_steps):
            r, u = run_step(ctx, pr, prec, 1e-10, torch, x, repl)
            check_step(name, r, f"step {k}")
            runs.append(r)
            log(f"step {name} {k}: {r['wall_ms']:.1f} ms (assemble {r['assemble_ms']:.2f} ms, "
                f"solve {r['solve_ms']:.1f} ms), it {r['niter']}, outer {r['n_outer']}, "
                f"rnorm {r['rnorm']:.2e}, spmv {r['spmv_ms']:.4f} ms x {r['spmv_calls']}")
        sols[name] = u.detach().cpu().numpy().copy()
        res[name] = runs

    if failed:
        print(json.dumps({"metric": METRIC, "error": "a timed or warm-up solve did not converge",
                          "failed": failed}), flush=True)
        sys.exit(3)
    mixed = res["mixed"]
    ms_step = float(np.mean([r["wall_ms"] for r in mixed]))
    value = pr.n_dof / (ms_step / 1e3)
    out = {"metric": METRIC, "value": value, "unit": "DOF/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32/f64", "data": "synthetic",
           "config": {"workload": (f"{n}^3 C3D8 unit cube, E lognormal(ln 1e9, 0.5) seed "
                                   f"{P.SEED}, nu 0.3, z=0 clamped (penalty 1e12), p=1e6 on "
                                   "z=1; assemble + Jacobi-PCG to 1e-10: FP32 element "
                                   "integration, values and inner PCG (tol 1e-5, FP64 "
                                   "residual replacement every 0.1 with the search "
                                   "direction kept), FP64 refinement residual, update and "
                                   "dots"),
                      "n_dof": pr.n_dof, "nnz": ctx.nnz, "n_elems": pr.n_elems,
                      "precision": "mixed"},
           "mixed": {"dof_per_s": value, "ms": ms_step,
                     "assemble_ms": float(np.mean([r["assemble_ms"] for r in mixed])),
                     "solve_ms": float(np.mean([r["solve_ms"] for r in mixed])),
                     "inner_iters": mixed[-1]["niter"], "outer_iters": mixed[-1]["n_outer"],
                     "fp64_residual_replacements": mixed[-1]["n_replace"],
                     "replace_rtol": args.replace_rtol, "inner_rtol": 1e-5,
                     "rnorm": mixed[-1]["rnorm"]}}
    if "mixed_restarted" in res:
        mr = res["mixed_restarted"]
        msr = float(np.mean([r["wall_ms"] for r in mr]))
        out["mixed_restarted"] = {
            "form": "each refinement step restarts the FP32 PCG (inner tol 1e-5, no replacement)",
            "dof_per_s": pr.n_dof / (msr / 1e3), "ms": msr, "steps": len(mr),
            "inner_iters": mr[-1]["niter"],
            "outer_iters": mr[-1]["n_outer"], "rnorm": mr[-1]["rnorm"]}
    if "fp64" in res:
        f64 = res["fp64"]
        ms64 = float(np.mean([r["wall_ms"] for r in f64]))
        out["fp64"] = {"dof_per_s": pr.n_dof / (ms64 / 1e3), "ms": ms64, "steps": len(f64),
                       "assemble_ms": float(np.mean([r["assemble_ms"] for r in f64])),
                       "solve_ms": float(np.mean([r["solve_ms"] for r in f64])),
                       "iters": f64[-1]["niter"], "rnorm": f64[-1]["rnorm"],
                       "spmv_ms": f64[-1]["spmv_ms"],
                       "spmv_gbs": spmv64_bytes(pr.n_dof, ctx.nnz) / (f64[-1]["spmv_ms"] * 1e6)}
        out["mixed_over_fp64"] = ms64 / ms_step
        out["errors"] = {f"{k}_vs_fp64_gpu_rel": float(np.linalg.norm(v - sols["fp64"]) /
                                                        np.linalg.norm(sols["fp64"]))
                         for k, v in sols.items() if k != "fp64"}
    # roofline of the dominant kernel (FP32 SpMV + p.Ap inside the mixed solve)
    spmv_ms = mixed[-1]["spmv_ms"]
    b32 = spmv32_bytes(pr.n_dof, ctx.nnz, ctx.n_nodes, ctx.nb_index_bytes,
                       int(np.unique(pr.bc_eq).size))
    k32 = SPMV32_KERNEL[ctx.nb_index_bytes].format(tile=ctx.sym_tile)
    ach = b32 / (spmv_ms * 1e6)
    copy_gbs = copy_bandwidth(torch, dev)
    out["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic(pr.n_dof, k32),
                       "copy_gbs": copy_gbs, "frac_of_copy": ach / copy_gbs,
                       "kernel": k32 + " (" + SPMV32_FORM[ctx.nb_index_bytes] + ")",
                       "bytes_per_launch": b32, "ms_per_launch": spmv_ms,
                       "traffic_source": "profiles/spmv_pcg128_hbm_traffic.md (rocprofv3 "
                                         "FETCH_SIZE / WRITE_SIZE, copy-calibrated)"}
    traffic = out["roofline"]["traffic"]
    if traffic:
        # what the kernel moves over HBM / the fabric per launch, at this run's
        # launch time (the symmetric view re-reads each block once more, partly
        # from L2, so this exceeds the algorithmic rate)
        out["roofline"]["traffic_gbs"] = traffic / (spmv_ms * 1e6)
        out["roofline"]["traffic_frac"] = traffic / (spmv_ms * 1e6) / HBM_PEAK_GBS
    # element integration + atomics-free scatter: VALU-bound; FLOPs counted in
    # the reference's dense B^T D B form (76,432 flop per element)
    # MI355X vector peaks: FP64 78.6 TFLOP/s; FP32 157.3 TFLOP/s with packed
    # FP32 (dual FMA per lane), which this build does not emit, so 78.6 is the
    # FP32 ceiling of the code as compiled
    asm = {"bound": "valu", "unit": "TFLOP/s", "flop_per_elem_ref_dense": 76432,
           "flop": 76432.0 * pr.n_elems, "peak_fp64_vector_tflops": 78.6,
           "peak_fp32_vector_tflops": 157.3, "fp32_ceiling_without_packed_tflops": 78.6,
           "fp32_ms": out["mixed"]["assemble_ms"]}
    asm["fp32_tflops"] = asm["flop"] / (asm["fp32_ms"] * 1e9)
    asm["fp32_frac"] = asm["fp32_tflops"] / asm["peak_fp32_vector_tflops"]
    asm["fp32_frac_of_unpacked"] = asm["fp32_tflops"] / asm["fp32_ceiling_without_packed_tflops"]
    if "fp64" in out:
        asm["fp64_ms"] = out["fp64"]["assemble_ms"]
        asm["fp64_tflops"] = asm["flop"] / (asm["fp64_ms"] * 1e9)
        asm["fp64_frac"] = asm["fp64_tflops"] / asm["peak_fp64_vector_tflops"]
    # the C-ABI also takes host arrays: the same assembly with E in host memory
    # (the PCIe-inclusive form of the step; the timed steps pass device data)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.assemble_k(pr.young_dev, pr.nu, precision=ufc_hip.PREC_MIXED)
    t1 = time.perf_counter()
    ctx.assemble_k(np.ascontiguousarray(pr.young), pr.nu, precision=ufc_hip.PREC_MIXED)
    t2 = time.perf_counter()
    asm["host_young_extra_ms"] = ((t2 - t1) - (t1 - t0)) * 1e3
    # element variants (UFC_VARIANT_*) through the same fused assembly: mean of
    # 3 launches after one warm-up, events around the assembly call
    variants = {}
    for vname, v in (("standard", 1), ("reduced", 2), ("selective", 3), ("incompat", 4),
                     ("gauss27", 6)):
        row = {}
        for pname, prec in (("fp32_ms", ufc_hip.PREC_MIXED), ("fp64_ms", ufc_hip.PREC_FP64)):
            ts = []
            for k in range(4):
                ctx.assemble_k(pr.young_dev, pr.nu, precision=prec, variant=v)
                if k:
                    ts.append(ctx.timing().assemble_ms)
            row[pname] = float(np.mean(ts))
        variants[vname] = row
        log(f"assembly variant {vname}: FP32 {row['fp32_ms']:.2f} ms, FP64 {row['fp64_ms']:.2f} ms")
    asm["variants"] = variants
    # y = K^T x through the mirrored blocks (FP64), beside y = K x (device data)
    ctx.assemble_k(pr.young_dev, pr.nu, precision=ufc_hip.PREC_FP64)
    xt = torch.randn(pr.n_dof, dtype=torch.float64, device=dev)
    yt = torch.empty_like(xt)
    tr = {}
    for name, fn in (("spmv_fp64_ms", lambda: ctx.spmv(xt, precision=ufc_hip.PREC_FP64, out=yt)),
                     ("spmv_t_fp64_ms", lambda: ctx.spmv_t(xt, out=yt))):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        tr[name] = (time.perf_counter() - t0) / 5 * 1e3
    tr["note"] = ("wall time per C-ABI call on device vectors (includes the call's own "
                  "synchronisation); K^T x reads each neighbour block transposed in place")
    log(f"K x {tr['spmv_fp64_ms']:.2f} ms, K^T x {tr['spmv_t_fp64_ms']:.2f} ms (FP64)")
    out["transpose"] = tr
    asm["variants_note"] = ("incompat includes its condensation-factor kernel (one wave per "
                            "element, 24x9 doubles per element in HBM)")
    ex = measured_assembly_flop(pr.n_elems)
    if ex is not None:
        asm["executed_flop"] = ex
        asm["executed_flop_source"] = ("profiles/assembly128_valu_flops.md (rocprofv3 "
                                       "SQ_INSTS_VALU_FLOPS_FP32/FP64 x 64 lanes)")
    out["assembly"] = asm
    if not args.no_views and "fp64" in out:
        log("storage vs precision")
        out["storage_vs_precision"] = storage_vs_precision(ctx, pr, torch, x, ufc_hip,
                                                           out["mixed"], out["fp64"], check_step,
                                                           args.replace_rtol)
        out["storage_vs_precision"]["mixed_symdiag"]["spmv_ms"] = out["roofline"]["ms_per_launch"]
        if failed:
            print(json.dumps({"metric": METRIC, "error": "a storage/precision solve did not "
                              "converge", "failed": failed}), flush=True)
            sys.exit(3)
    ctx.close()
    del pr.young_dev

    if not args.no_renumbered:
        out["renumbered"] = renumbered_run(args, torch, dev, ufc_hip, P, n, sols)

    if not args.no_shard:
        out["shard"] = shard_run(args, torch, dev, ufc_hip, P, n, check_step)
        if failed:
            print(json.dumps({"metric": METRIC, "error": "a shard solve did not converge",
                              "failed": failed}), flush=True)
            sys.exit(3)

    if not args.no_adjoint:
        na = args.adjoint_size
        log(f"adjoint {na}^3 (torch.autograd.Function over the C-ABI)")
        from ufc_hip.autograd import StaticModel, compliance
        pa = P.simp_cube(na)
        model = StaticModel.from_problem(pa, precision=ufc_hip.PREC_MIXED, rtol=1e-10)
        rho0 = torch.as_tensor(pa.extra["rho"], device=dev)
        for w in range(args.warmup + 1):
            rho = rho0.clone().requires_grad_(True)
            torch.cuda.synchronize()
            ta = time.perf_counter()
            c, u = compliance(model, 1e-3 + rho ** 3 * 1e9)
            torch.cuda.synchronize()
            tb = time.perf_counter()
            c.backward()
            torch.cuda.synchronize()
            tc = time.perf_counter()
        fwd_ms = (tb - ta) * 1e3
        adj_ms = (tc - tb) * 1e3
        # the GPU's E field, forward solution and mixed gradient of the last
        # step, for the oracle's checks at full size
        young_h = (1e-3 + rho0 ** 3 * 1e9).cpu().numpy()
        u_h = u.detach().cpu().numpy()
        g_mixed = rho.grad.detach().cpu().numpy()
        g_fp64 = None
        out["adjoint"] = {"workload": (f"{na}^3 SIMP cube, rho~U[0.3,1] seed {P.SEED + 1}, "
                                       "E = 1e-3 + rho^3 1e9; c = f.u; dc/drho by "
                                       "torch.autograd (adjoint PCG + element kernel)"),
                          "n_dof": pa.n_dof, "precision": "mixed", "forward_ms": fwd_ms,
                          "grad_ms": adj_ms, "grad_over_forward": adj_ms / fwd_ms,
                          "fwd_iters": model.last_forward.niter,
                          "adj_iters": model.last_adjoint.niter,
                          "compliance": float(c)}
        log(f"adjoint: forward {fwd_ms:.1f} ms, grad {adj_ms:.1f} ms")
        if args.fd_check:
            # central differences h = 1e-6 on seeded elements, FP64 solves to 1e-14
            log(f"finite-difference check on {args.fd_check} elements")
            m64 = StaticModel.from_problem(pa, precision=ufc_hip.PREC_FP64, rtol=1e-14,
                                           adjoint_rtol=1e-14)
            rho = rho0.clone().requires_grad_(True)
            c64, _ = compliance(m64, 1e-3 + rho ** 3 * 1e9)
            c64.backward()
            g = rho.grad.detach().cpu().numpy()
            g_fp64 = g
            gm = rho0.clone().requires_grad_(True)
            cm, _ = compliance(model, 1e-3 + gm ** 3 * 1e9)
            cm.backward()
            gmx = gm.grad.detach().cpu().numpy()
            picks = np.random.default_rng(P.SEED + 3).choice(pa.n_elems, args.fd_check,
                                                             replace=False)
            errs = []
            with torch.no_grad():
                for e in picks:
                    rp_, rm_ = rho0.clone(), rho0.clone()
                    rp_[e] += 1e-6
                    rm_[e] -= 1e-6
                    cp, _ = compliance(m64, 1e-3 + rp_ ** 3 * 1e9)
                    cn, _ = compliance(m64, 1e-3 + rm_ ** 3 * 1e9)
                    fd = (float(cp) - float(cn)) / 2e-6
                    errs.append(abs(fd - g[e]) / abs(g[e]))
                    if len(errs) % 4 == 0:
                        log(f"fd {len(errs)}/{len(picks)}: max rel err {max(errs):.2e}")
            out["adjoint"]["fd_check"] = {
                "elements": int(args.fd_check), "h": 1e-6, "solve_rtol": 1e-14,
                "max_rel_err_fp64_adjoint_vs_fd": float(max(errs)),
                "median_rel_err": float(np.median(errs)),
                "mixed_vs_fp64_grad_rel": float(np.abs(gmx - g).max() / np.abs(g).max())}
            log(f"fd check: max rel err {max(errs):.2e}")
            m64.close()
        model.close()
        if not args.no_cpu:
            out["adjoint"]["cpu_baseline"] = cpu_adjoint_baseline(
                pa, young_h, u_h, {"mixed": g_mixed, "fp64": g_fp64})
            out["adjoint"]["errors"] = out["adjoint"]["cpu_baseline"].pop("errors")
            log(f"cpu adjoint: {out['adjoint']['cpu_baseline']}")

    if not args.no_j2:
        run_j2(args, out, P, ufc_hip, torch)

    if not args.no_cpu:
        log("cpu baseline + full-size errors against the oracle")
        out["cpu_baseline"], cerr = cpu_baseline(pr, sols)
        log(f"cpu {out['cpu_baseline']['value']:.3e} DOF/s; errors {cerr}")
        so = cpu_second_opinion()
        if so is not None:
            out["cpu_baseline"]["small_config_second_opinion"] = so
            log(f"cpu second opinion (cantilever): {so}")
        out.setdefault("errors", {}).update(cerr)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
