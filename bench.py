#!/usr/bin/env python3
"""Benchmark of the finite-element hot path on one MI355X.

Workload (default): 3D linear elasticity on the unit cube, 64^3 hexahedra
split into 6 P1 tetrahedra (1,572,864 elements, 274,625 nodes, 811,200 free
DOFs after clamping x=0), lognormal Young's modulus (sigma 1, seed 0,
64 quantised bins), Poisson ratio 0.3, 4-point Gauss, unit traction on x=1.
One step = Generate[A] (device assembly) + Solve[A] with Sparskit-style CG on
the diagonally scaled system ("Jacobi", Scaling 1) to 1e-10, FP64.

Prints one JSON line.  `value` = seconds per step (assembly + solve to
1e-10 relative residual); also reported: assembly elements/s, the SpMV
roofline (algorithmic bytes / average kernel time vs 8 TB/s), the learned
preconditioner against plain ILU(0)-GMRES on an unseen 256^2 field (with the
device training-step rate), the 512-instance batch (solves/s, latency
percentiles) and a bounded CPU baseline (the oracle, single-threaded).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("s to 1e-10 residual (learned vs plain precond); assembly elems/s; SpMV %HBM")
HBM_PEAK_GBS = 8000.0
# HBM bytes per SpMV launch on the CG operator of elasticity3d_n (RCMK order,
# 16-bit column offsets), from two rocprofv3 --pmc runs of `bench.py
# --roofline-only` (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, factors from a
# 512 MiB device copy): 403.5 MB read + 6.5 MB written.
PMC_SPMV_TRAFFIC = {64: 409_978_000}
# HBM bytes per assembly of elasticity3d_n (k_geometry + k_elem_nodal +
# k_slot_sum + k_rhs_sum), from the same two counter runs: 6.29 GB read +
# 2.36 GB written.
PMC_ASSEMBLY_TRAFFIC = {64: 8_645_000_000}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(N: int, nnz: int, s_val=8, s_x=8) -> int:
    """Algorithmic bytes of one CSR SpMV: values + column indices once, row
    pointers once, x read once, y written once."""
    return nnz * (s_val + 4) + (N + 1) * 4 + N * (s_x + s_x)


def ilu_apply_bytes(N: int, nnz: int, s_val: int = 4) -> int:
    """Algorithmic bytes of one ILU(0) application (both sweeps): factor
    entries (value + 4-byte column) once, row pointers/pivots, the input
    read once, the result written once per sweep."""
    return (nnz - N) * (s_val + 4) + N * (4 + 8) + 2 * N * 8


def assembly_bytes(pr, N: int, nnz: int) -> int:
    """Compulsory bytes of Generate[A]: connectivity + region per element,
    node coordinates once, CSR values and the RHS written once."""
    E = int((pr.etype == 4).sum() + (pr.etype == 2).sum())
    return E * (4 * 4 + 4) + pr.nnodes * 3 * 8 + nnz * 8 + N * 8


def cpu_baseline(pr, its_gpu: int, cg_sample: int):
    """Oracle (the C restatement; 1 thread: the reference's order, more
    threads: threaded element loop, amux and vector loops) on a bounded sample
    of the same workload: full assembly of the same mesh, then `cg_sample` CG
    iterations on the same system; the solve time is scaled to the GPU's
    iteration count."""
    from oracle import pyoracle as O
    t0 = time.perf_counter()
    ia, ja, a, b, tab = O.assemble(pr)
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    x, info, _ = O.solve(ia, ja, a, b, algorithm=1, preconditioner=0, scaling=1,
                         maxits=cg_sample, renumbering=1)
    t_cg = time.perf_counter() - t0
    per_it = t_cg / max(info["its"], 1)
    value = t_asm + per_it * its_gpu
    return {"value": round(value, 3), "unit": "s", "kind": "port",
            "sample": (f"oracle Generate of the full mesh ({t_asm:.2f} s) + {info['its']} "
                       f"CG+Jacobi iterations ({per_it * 1e3:.1f} ms/it), scaled to the "
                       f"device's {its_gpu} iterations")}


def cpu_threads() -> int:
    """Host threads available to this process (affinity mask, capped by
    OMP_NUM_THREADS when set)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_pair(fn):
    """Run fn() with the oracle on 1 thread and on every available thread;
    returns [result_1, result_n], each with "cores" and "kind" set."""
    from oracle import pyoracle as O
    out = []
    for T in sorted({1, cpu_threads()}):
        O.set_threads(T)
        try:
            r = fn()
        finally:
            O.set_threads(1)
        r.update(cores=T, kind="port")
        out.append(r)
    return out


def cpu_gmres(pr, its_gpu: int = 0, sample: int = 0, algorithm: int = 8, precond: int = 6, krylov: int = 30):
    """The oracle's solve_matrix (RCMK, the reference's ordering) on problem
    pr: the whole solve, or with `sample` > 0 that many iterations after the
    set-up, the iteration time scaled to the device's `its_gpu` iterations."""
    from oracle import pyoracle as O

    def run():
        ia, ja, a, b, tab = O.assemble(pr)
        t0 = time.perf_counter()
        x, info, _ = O.solve(ia, ja, a, b, algorithm=algorithm, preconditioner=precond, krylov=krylov,
                             maxits=sample if sample else 100000, renumbering=1,
                             scaling=1 if precond == 0 else 0)
        t = time.perf_counter() - t0
        if sample:
            # set-up alone (renumbering, factorisation): one iteration
            t1 = time.perf_counter()
            O.solve(ia, ja, a, b, algorithm=algorithm, preconditioner=precond, krylov=krylov, maxits=1,
                    renumbering=1, scaling=1 if precond == 0 else 0)
            setup = time.perf_counter() - t1
            per = max(t - setup, 0.0) / max(info["its"] - 1, 1)
            return {"value_ms": round(1e3 * (setup + per * its_gpu), 1), "its": its_gpu,
                    "sample": f"{info['its']} iterations timed ({1e3 * per:.2f} ms/it) + set-up "
                              f"{1e3 * setup:.0f} ms, scaled to {its_gpu} iterations"}
        return {"value_ms": round(1e3 * t, 1), "its": info["its"], "status": info["status"],
                "sample": "whole solve"}
    return cpu_pair(run)


def unet_flops_per_pixel() -> float:
    """Algorithmic flops of one U-Net forward pass per full-resolution pixel:
    2*9*cin*cout per pixel of each 3x3 convolution (true channel counts; a
    level-l convolution runs on 4^-l of the pixels) + the 1x1 output layer."""
    spec = [(2, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2),
            (64, 128, 3), (128, 128, 3), (192, 64, 2), (64, 64, 2), (96, 32, 1), (32, 32, 1),
            (48, 16, 0), (16, 16, 0)]
    return sum(2 * 9 * ci * co / 4 ** lvl for ci, co, lvl in spec) + 2 * 16


def learned_section(n: int = 256, seed: int = 90000):
    """Learned two-stage FGMRES (shipped weights) against plain ILU(0)-GMRES on
    an unseen n x n checkerboard field, and the device training-step rate."""
    from getdp_amd import learned as L
    from getdp_amd import linalg as LA
    from pathlib import Path
    w = Path(ROOT) / "getdp_amd" / "weights" / "unet_poisson2d.npz"
    with np.load(w, allow_pickle=False) as z:
        params = z["params"]
    row = L.evaluate(params, n, [seed], log=log)[0]
    # C0's plain method: one ILU(0) application in the reference's RCMK order
    pr = L.random_field_problem(n, seed)
    fe, A, b = LA.assemble(pr)
    s0 = LA.Solver(Algorithm=8, Preconditioner=6, Krylov_Size=30, Nb_Iter_Max=1, Stopping_Test=1e-10,
                   Renumbering_Technique=1)
    i0 = LA.solve(A, b, s0, LA.Vector(fe.N))
    c0_apply = LA.solver_prec_apply(s0, b, LA.Vector(fe.N), reps=20)
    c0_bytes = ilu_apply_bytes(fe.N, fe.nnz)
    s0.close()
    # training-step throughput at batch 32, 128^2 (forward + backward + Adam)
    net = L.UNet(128, 128, batch=32, seed=0)
    rng = np.random.default_rng(0)
    r = rng.standard_normal((32, 128, 128)).astype(np.float32)
    f = np.zeros_like(r)
    e = rng.standard_normal((32, 128, 128)).astype(np.float32)
    e /= np.linalg.norm(e.reshape(32, -1), axis=1)[:, None, None]
    mask = np.ones((128, 128), np.float32)
    for _ in range(3):
        net.train_step(r, f, e, mask, 1e-4)
    LA.lib().gdp_device_synchronize()
    t0 = time.perf_counter()
    k = 20
    for _ in range(k):
        net.train_step(r, f, e, mask, 1e-4)
    LA.lib().gdp_device_synchronize()
    step = (time.perf_counter() - t0) / k
    flops = 3 * unet_flops_per_pixel() * 32 * 128 * 128  # forward + input grad + weight grad
    return {
        "workload": f"poisson2d_{n} checkerboard seed {seed} (unseen), 1e-10",
        "learned_ms": min(row["learned_rcmk"]["ms"], row["learned_multicolour"]["ms"]),
        "learned_its": row["learned_multicolour"]["its"] if row["learned_multicolour"]["ms"] <= row["learned_rcmk"]["ms"] else row["learned_rcmk"]["its"],
        "plain_rcmk_ms": row["plain_rcmk"]["ms"], "plain_rcmk_its": row["plain_rcmk"]["its"],
        "plain_multicolour_ms": row["plain_multicolour"]["ms"],
        "speedup_vs_plain_rcmk": row["speedup_vs_plain_rcmk"],
        "speedup_vs_best_plain": row["speedup_vs_best_plain"],
        "rel_res": max(row["learned_rcmk"]["rel_res"], row["learned_multicolour"]["rel_res"]),
        "c0_ilu_apply_ms": round(c0_apply, 4), "c0_ilu_levels": i0["ilu_levels"],
        "c0_ilu_apply_gbs": round(c0_bytes / (c0_apply * 1e-3) / 1e9, 1),
        "train_step_ms_b32_128": round(step * 1e3, 2),
        "train_tflops": round(flops / step / 1e12, 2),
        "train_mfma_frac_bf16": round(flops / step / 2.5e15, 4),
    }


def learned3d_section(n: int = 64, seed: int = 90002):
    """The learned preconditioner on the 3D problem (poisson3d_n: Kuhn
    tetrahedra on an n^3 grid, unseen checkerboard coefficient field), shipped
    weights: two-stage FGMRES in FP64 and in mixed precision (FP32 inner
    FGMRES, FP64 refinement) against the plain GPU ILU(0)-GMRES(30) in the
    reference's RCMK ordering (and in multicolour ordering), all to 1e-10."""
    from getdp_amd import learned3d as L3
    from pathlib import Path
    w = Path(ROOT) / "getdp_amd" / "weights" / "unet_poisson3d.npz"
    with np.load(w, allow_pickle=False) as z:
        params = z["params"]
        lsq = bool(z["lsq"]) if "lsq" in z.files else False
    row = L3.evaluate(params, n, [seed], log=log, precisions=(0, 1), lsq=lsq)[0]
    pick = lambda keys: min(keys, key=lambda k: row[k]["ms"] if row[k]["rel_res"] <= 1.5e-10 else 1e30)
    b64 = pick(["learned_rcmk", "learned_multicolour"])
    bmx = pick(["learned_rcmk_mixed", "learned_multicolour_mixed"])
    ref = row["plain_rcmk"]["ms"]
    return {
        "workload": f"poisson3d_{n} checkerboard seed {seed} (unseen): {row['dofs']} DOFs, 1e-10",
        "plain_rcmk": row["plain_rcmk"], "plain_multicolour": row["plain_multicolour"],
        "learned_fp64": dict(row[b64], variant=b64), "learned_mixed": dict(row[bmx], variant=bmx),
        "speedup_vs_plain_rcmk": round(ref / row[b64]["ms"], 2),
        "speedup_mixed_vs_plain_rcmk": round(ref / row[bmx]["ms"], 2),
        "speedup_vs_best_plain": row["speedup_vs_best_plain"],
        "target_speedup": 5.0,
    }


def c2_section(n: int = 48, its64: int = 1500):
    """Curl-curl with Whitney edges (1e-6 mass, 1:1000 reluctivity): device
    assembly, then ILU(0)-GMRES(30) (multicolour ordering) in FP64 for a fixed
    iteration budget, and the nested mixed-precision solve (FP64 FGMRES around
    20 FP32 GMRES steps) stopped at the residual the FP64 run reached; reports
    both times and the difference of the two solutions."""
    from getdp_amd import linalg as LA
    from getdp_amd import problems as P
    pr = P.curlcurl3d(n)
    fe = LA.FESystem(pr)
    A = LA.Matrix(fe.N)
    b = LA.Vector(fe.N)
    fe.generate(A, b)
    fe.generate(A, b)
    asm_ms = fe.timing()["ms_total"]
    E = int((pr.etype == P.TET).sum())
    base = dict(Algorithm=8, Preconditioner=6, Krylov_Size=30, ILU_Ordering=1, Orthogonalization=1)
    s64 = LA.Solver(Nb_Iter_Max=its64, Stopping_Test=1e-10, **base)
    x64 = LA.Vector(fe.N)
    i64 = LA.solve(A, b, s64, x64)
    target = max(1e-10, i64["rel_res"])
    kw = dict(base)
    kw.update(Algorithm=9, Precision=2, Inner_Krylov_Size=20)
    s32 = LA.Solver(Nb_Iter_Max=300, Stopping_Test=target, **kw)
    x32 = LA.Vector(fe.N)
    i32 = LA.solve(A, b, s32, x32)
    a, c = x64.get(), x32.get()
    # the reference's own FP64 solution (CG/Jacobi to 1e-10 in the reference's
    # Fortran, sampled every 97th entry: tests/golden/c2_curlcurl48_cg.npz)
    vs_oracle = None
    gold = os.path.join(ROOT, "tests", "golden", "c2_curlcurl48_cg.npz")
    if n == 48 and os.path.exists(gold):
        with np.load(gold, allow_pickle=False) as z:
            st, xr, sc = int(z["stride"]), z["x_sampled"], float(z["norm_inf"])
        vs_oracle = lambda v: float(np.abs(v[::st] - xr).max() / sc)
    # the SPD operator also takes Sparskit's CG on the diagonally scaled system,
    # which does reach 1e-10 (Stopping_Test on the scaled residual, true residual reported)
    scg = LA.Solver(Algorithm=1, Preconditioner=0, Scaling=1, Renumbering_Technique=1,
                    Nb_Iter_Max=200000, Stopping_Test=1e-10)
    xcg = LA.Vector(fe.N)
    icg = LA.solve(A, b, scg, xcg)
    # where the GMRES and CG solutions differ: the curl-curl operator alone
    # (mass 0) applied to their difference, relative to it applied to the CG
    # solution; near zero means the difference is a discrete gradient, which
    # only the 1e-6 mass term sees
    pr0 = P.curlcurl3d(n, mass=0.0)
    fe0 = LA.FESystem(pr0)
    K0, b0 = LA.Matrix(fe0.N), LA.Vector(fe0.N)
    fe0.generate(K0, b0)
    xc = xcg.get()
    d, y = LA.Vector(fe.N, a - xc), LA.Vector(fe.N)
    LA.prod_matrix_vector(K0, d, y)
    kd = np.linalg.norm(y.get())
    LA.prod_matrix_vector(K0, LA.Vector(fe.N, xc), y)
    curl_frac = float(kd / max(np.linalg.norm(y.get()), 1e-300))
    return {
        "workload": f"curlcurl3d_{n}: {E} tets, {fe.N} edge DOFs, nnz {fe.nnz}",
        "assembly_ms": round(asm_ms, 3), "assembly_elems_per_s": round(E / (asm_ms * 1e-3), 1),
        "fp64": {"its": i64["its"], "rel_res": i64["rel_res"], "ms": round(i64["t_solve_ms"] + i64["t_setup_ms"], 1),
                 "error_vs_oracle": vs_oracle(a) if vs_oracle else None},
        "mixed_nested": {"outer_its": i32["its"], "inner_fp32_steps": 20 * i32["its"],
                         "rel_res": i32["rel_res"], "ms": round(i32["t_solve_ms"] + i32["t_setup_ms"], 1),
                         "error_vs_oracle": vs_oracle(c) if vs_oracle else None},
        "solution_rel_diff": float(np.abs(a - c).max() / max(np.abs(a).max(), 1e-300)),
        "fp64_cg_jacobi": {"its": icg["its"], "status": icg["status"], "rel_res": icg["rel_res"],
                           "ms": round(icg["t_solve_ms"] + icg["t_setup_ms"], 1),
                           "error_vs_oracle": vs_oracle(xcg.get()) if vs_oracle else None,
                           "gmres_solution_rel_diff": float(np.abs(a - xc).max() / max(np.abs(xc).max(), 1e-300)),
                           "gmres_diff_curl_fraction": curl_frac},
        "note": "ILU(0)-GMRES(30) stagnates near 6e-7 on this operator in FP64 as well; both GMRES runs are "
                "compared at the residual the FP64 run reaches; CG on the scaled system reaches 1e-10; "
                "error_vs_oracle = max |x - x_ref| / max |x_ref| over every 97th entry of the reference's "
                "FP64 CG/Jacobi solution; gmres_diff_curl_fraction = |K_curl (x_gmres - x_cg)| / |K_curl x_cg| "
                "with K_curl the operator without its 1e-6 mass term: the GMRES error sits in the gradient "
                "fields the curl does not see",
    }


def default_solver_section(n: int = 512):
    """GetDP's default linear solver (Preconditioner 2 = ILUTP, Nb_Fill 20,
    Permutation_Tolerance 0.05, no dropping; GMRES(40); RCMK) on the 2D
    Poisson problem poisson2d_n, once, outside the timed steps: set-up (RCMK,
    host ILUTP factorisation, upload of the factors) and the device GMRES."""
    from getdp_amd import linalg as LA
    from getdp_amd import problems as P
    pr = P.poisson2d(n)
    fe = LA.FESystem(pr)
    A, b, x = LA.Matrix(fe.N), LA.Vector(fe.N), LA.Vector(fe.N)
    fe.generate(A, b)
    s = LA.Solver()  # every parameter at its default
    info = LA.solve(A, b, s, x)
    again = LA.solve_again(A, b, s, x)  # factors kept: the GMRES phase alone
    apply_ms = LA.solver_prec_apply(s, b, LA.Vector(fe.N), reps=5)
    res = {"workload": f"poisson2d_{n}: {fe.N} DOFs, nnz {fe.nnz}", "its": info["its"],
           "status": info["status"], "rel_res": info["rel_res"],
           "setup_ms": round(info["t_setup_ms"], 1), "solve_ms": round(info["t_solve_ms"], 1),
           "setup_factor_ms": round(info["t_factor_ms"], 1), "setup_load_ms": round(info["t_load_ms"], 1),
           "setup_other_ms": round(info["t_setup_ms"] - info["t_factor_ms"] - info["t_load_ms"], 1),
           "solve_again_ms": round(again["t_solve_ms"], 1), "ilu_levels": info["ilu_levels"],
           "ilu_apply_ms": round(apply_ms, 3),
           "total_ms": round(info["t_setup_ms"] + info["t_solve_ms"], 1)}
    log(f"default solver (ILUTP(20)-GMRES(40)) poisson2d_{n}: its={info['its']} status={info['status']} "
        f"setup {info['t_setup_ms']:.0f} ms solve {info['t_solve_ms']:.0f} ms")
    s.close()
    return res, pr


def default_solver_cpu(pr):
    """The oracle's solve_matrix with the same defaults on the same system:
    the reference's CPU path for this solver (the ILUTP factorisation and the
    triangular solves are sequential at any thread count)."""
    from oracle import pyoracle as O
    ia, ja, a, b, tab = O.assemble(pr)

    def run():
        t0 = time.perf_counter()
        x, info, _ = O.solve(ia, ja, a, b, algorithm=8, preconditioner=2, krylov=40, maxits=1000,
                             renumbering=1)
        return {"value_ms": round(1e3 * (time.perf_counter() - t0), 1), "its": info["its"],
                "status": info["status"]}
    return cpu_pair(run)


def batch_section(count: int = 512):
    from getdp_amd import batch as BT
    return BT.bench(count, reps=1, log=log)


def batch_cpu(count: int = 512, stride: int = 32):
    """The oracle on every `stride`-th of the same instances (assembly and
    ILU(0)-GMRES(30) to 1e-10 each): on 1 thread in sequence (time scaled by
    `stride`), and on a pool of one worker per host thread (each solve
    sequential), where the estimate is the sum of the instance times measured
    in the pool, scaled by `stride`, over the thread count -- a perfectly
    balanced pool over all instances, which favours the CPU."""
    from concurrent.futures import ThreadPoolExecutor
    from getdp_amd import batch as BT
    from oracle import pyoracle as O
    probs = BT.make_instances(count, seed=0)[::stride]

    def one(p):
        t0 = time.perf_counter()
        ia, ja, a, b, tab = O.assemble(p)
        x, info, _ = O.solve(ia, ja, a, b, algorithm=8, preconditioner=6, krylov=30, maxits=100000,
                             renumbering=1)
        return info["status"], time.perf_counter() - t0

    out = []
    O.set_threads(1)
    for T in sorted({1, cpu_threads()}):
        if T == 1:
            res = [one(p) for p in probs]
        else:
            with ThreadPoolExecutor(T) as ex:
                res = list(ex.map(one, probs))
        t = sum(r[1] for r in res) * stride / T
        out.append({"value_ms": round(1e3 * t, 1), "solves_per_s": round(count / t, 2), "cores": T,
                    "kind": "port", "converged_in_sample": int(sum(1 for r in res if r[0] == 0)),
                    "sample": (f"{len(probs)} of the {count} instances (every {stride}th); "
                               + ("sequential, scaled" if T == 1 else
                                  f"instance times measured on a {T}-worker pool, scaled, over {T} workers"))})
    return out


def roofline_only(n: int, reps: int):
    """For rocprofv3 --pmc runs: the C1 matrix's SpMV `reps` times, then a
    plain device copy of 512 MiB (calibration of FETCH_SIZE/WRITE_SIZE)."""
    import torch
    from getdp_amd import linalg as LA
    from getdp_amd import problems as P
    pr = P.elasticity3d(n)
    LA.init(0)
    fe = LA.FESystem(pr)
    A = LA.Matrix(fe.N)
    b = LA.Vector(fe.N)
    fe.generate(A, b)
    x = LA.Vector(fe.N)
    # one CG iteration builds the solver's renumbered SELL copy (the CG operator)
    s = LA.Solver(Algorithm=1, Preconditioner=0, Scaling=1, Nb_Iter_Max=1, Stopping_Test=1e-10,
                  Renumbering_Technique=1)
    LA.solve(A, b, s, x)
    ms = LA.bench_solver_spmv(s, reps=reps)["ms"]
    src = torch.empty(64 * 2 ** 20, dtype=torch.float64, device="cuda:0").uniform_()
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"spmv_ms": ms, "spmv_algorithmic_bytes": spmv_bytes(fe.N, fe.nnz),
                      "copy_bytes_each_way": src.numel() * 8, "N": fe.N, "nnz": fe.nnz}))


def ilu_profile(n: int, reps: int = 20):
    """For rocprofv3 runs: C1's CG/ILU(0) set-up on elasticity3d_n (Scaling 1,
    RCMK), then `reps` ILU(0) applications alone."""
    from getdp_amd import linalg as LA
    from getdp_amd import problems as P
    LA.init(0)
    pr = P.elasticity3d(n)
    fe, A, b = LA.assemble(pr)
    s = LA.Solver(Algorithm=1, Preconditioner=6, Scaling=1, Nb_Iter_Max=1, Stopping_Test=1e-10,
                  Renumbering_Technique=1)
    info = LA.solve(A, b, s, LA.Vector(fe.N))
    ms = LA.solver_prec_apply(s, b, LA.Vector(fe.N), reps=reps)
    nb = ilu_apply_bytes(fe.N, fe.nnz)
    print(json.dumps({"n": n, "N": fe.N, "nnz": fe.nnz, "ilu_levels": info["ilu_levels"],
                      "ilu_sweep": info["ilu_sweep"], "apply_ms": round(ms, 4), "algorithmic_bytes": nb,
                      "gbs": round(nb / (ms * 1e-3) / 1e9, 1),
                      "us_per_level": round(1e3 * ms / (2 * info["ilu_levels"]), 3)}))


def unet_profile(mode: str, reps: int = 5):
    """For rocprofv3 runs: U-Net launches alone.  fwd3d: forward passes of the
    3D network on a 64^3 grid (batch 1, the preconditioner's shape); train3d:
    training steps at 32^3 x 16; train2d: training steps at 128^2 x 32 (the
    training line of the learned section)."""
    from getdp_amd import learned as L
    from getdp_amd import linalg as LA
    LA.init(0)
    rng = np.random.default_rng(0)
    if mode == "fwd3d":
        n, B, D = 64, 1, 64
    elif mode == "train3d":
        n, B, D = 32, 16, 32
    else:
        n, B, D = 128, 32, 1
    shape = (B, D, n, n) if D > 1 else (B, n, n)
    net = L.UNet(n, n, batch=B, seed=0, D=D)
    r = rng.standard_normal(shape).astype(np.float32)
    f = np.zeros_like(r)
    e = rng.standard_normal(shape).astype(np.float32)
    mask = np.ones(shape[1:], np.float32)
    run = (lambda: net.forward(r, f)) if mode == "fwd3d" else (lambda: net.train_step(r, f, e, mask, 1e-4))
    for _ in range(3):
        run()
    LA.lib().gdp_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    LA.lib().gdp_device_synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / reps
    px = B * n * n * (D if D > 1 else 1)
    fpp = unet_flops_per_voxel_3d() if D > 1 else unet_flops_per_pixel()
    flops = (1 if mode == "fwd3d" else 3) * fpp * px
    print(json.dumps({"mode": mode, "ms": round(ms, 3), "algorithmic_gflop": round(flops / 1e9, 2),
                      "tflops": round(flops / ms / 1e9, 2)}))


def unet_flops_per_voxel_3d() -> float:
    """As unet_flops_per_pixel for the 3D network: 27 taps, a level-l
    convolution on 8^-l of the voxels."""
    spec = [(2, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2),
            (64, 128, 3), (128, 128, 3), (192, 64, 2), (64, 64, 2), (96, 32, 1), (32, 32, 1),
            (48, 16, 0), (16, 16, 0)]
    return sum(2 * 27 * ci * co / 8 ** lvl for ci, co, lvl in spec) + 2 * 16


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="cells per cube side")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cg-sample", type=int, default=60)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain line")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-instance batch line")
    ap.add_argument("--no-c2", action="store_true", help="skip the curl-curl mixed-precision line")
    ap.add_argument("--no-default-solver", action="store_true",
                    help="skip the ILUTP(20)-GMRES(40) line (GetDP's default solver)")
    ap.add_argument("--default-solver-only", type=int, default=0, metavar="N",
                    help="only GetDP's default solver (ILUTP(20)-GMRES(40)) on poisson2d_N")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the SpMV launches and a calibration copy (for PMC runs)")
    ap.add_argument("--unet-profile", choices=["fwd3d", "train3d", "train2d"],
                    help="only U-Net launches (for rocprofv3 runs)")
    ap.add_argument("--ilu-profile", type=int, default=0, metavar="N",
                    help="only ILU(0) applications of CG/ILU(0) on elasticity3d_N (for rocprofv3 runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.roofline_only:
        roofline_only(args.n, 20)
        return
    if args.unet_profile:
        unet_profile(args.unet_profile)
        return
    if args.ilu_profile:
        ilu_profile(args.ilu_profile)
        return
    if args.default_solver_only:
        from getdp_amd import linalg as LA
        LA.init(0)
        res, _ = default_solver_section(args.default_solver_only)
        print(json.dumps(res))
        return

    from getdp_amd import linalg as LA
    from getdp_amd import problems as P

    t0 = time.perf_counter()
    pr = P.elasticity3d(args.n)
    log(f"mesh: {pr.nelem} elements, {pr.nnodes} nodes ({time.perf_counter() - t0:.1f} s)")
    LA.init(0)
    t0 = time.perf_counter()
    fe = LA.FESystem(pr)
    log(f"pre-processing (DOF numbering + pattern): N={fe.N} nnz={fe.nnz} "
        f"({time.perf_counter() - t0:.1f} s)")
    A = LA.Matrix(fe.N)
    b = LA.Vector(fe.N)
    x = LA.Vector(fe.N)
    solver = LA.Solver(Algorithm=1, Preconditioner=0, Scaling=1, Nb_Iter_Max=100000,
                       Stopping_Test=1e-10, Renumbering_Technique=1)

    def step():
        fe.generate(A, b)
        return LA.solve(A, b, solver, x)

    for w in range(args.warmup):
        info = step()
        log(f"warmup {w}: its={info['its']} rel_res={info['rel_res']:.2e}")
    LA.lib().gdp_device_synchronize()
    t_steps, asm_ms, infos = [], [], []
    for k in range(args.steps):
        t0 = time.perf_counter()
        info = step()
        LA.lib().gdp_device_synchronize()
        t_steps.append(time.perf_counter() - t0)
        asm_ms.append(fe.timing()["ms_total"])
        infos.append(info)
        log(f"step {k}: {t_steps[-1]:.3f} s  assembly {asm_ms[-1]:.2f} ms  its={info['its']} "
            f"rel_res={info['rel_res']:.2e} status={info['status']}")
    ms_per_step = 1e3 * float(np.mean(t_steps))
    E = int((pr.etype == 4).sum())
    asm_s = float(np.mean(asm_ms)) / 1e3

    # SpMV roofline on the CG's operator: the solver's RCMK-renumbered, scaled
    # SELL copy (the SpMV is the dominant kernel of the solve)
    sp = LA.bench_solver_spmv(solver, reps=50)
    spmv_ms = sp["ms"]
    sb = spmv_bytes(fe.N, fe.nnz)
    achieved = sb / (spmv_ms * 1e-3) / 1e9
    cols = "16-bit column offsets" if sp["compressed"] else "32-bit columns"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": PMC_SPMV_TRAFFIC.get(args.n),
                "kernel": f"k_sell_spmv<2> on the CG operator (RCMK order, SELL-64, FP64, {cols})",
                "algorithmic_bytes": sb, "avg_ms": round(spmv_ms, 5)}
    xs = LA.Vector(fe.N, np.random.default_rng(0).standard_normal(fe.N))
    ys = LA.Vector(fe.N)
    spmv_natural_ms = LA.bench_spmv(A, xs, ys, reps=50)

    # post-processing of the solution (PostOperation Print[...]): {d u_x} per
    # element and the energy integral, device time of the evaluation kernels
    vol = sorted(set(int(r) for r in pr.region[pr.etype == P.TET]))
    pd = fe.post(x, LA.PQ_D, vol, space=0)
    pd = fe.post(x, LA.PQ_D, vol, space=0)
    t0_ = pr.terms[0]
    pe = fe.post(x, LA.PQ_INT_COEF_D2, vol, space=0, coef=np.ones((int(pr.region.max()) + 1, 1)),
                 npoints=t0_.npoints)
    nsel = pd["nsel"]
    post_bytes = nsel * (4 * 4 + 3 * 4 + 12 * 8) + pr.nnodes * (3 * 8 + 4 + 8)
    post = {"elements": nsel, "d_kernel_ms": round(pd["kernel_ms"], 4),
            "d_elements_per_s": round(nsel / (pd["kernel_ms"] * 1e-3), 1),
            "d_algorithmic_bytes": post_bytes,
            "d_gbs": round(post_bytes / (pd["kernel_ms"] * 1e-3) / 1e9, 1),
            "energy_kernel_ms": round(pe["kernel_ms"], 4)}
    log(f"post: {{d u}} on {nsel} elements {pd['kernel_ms']:.3f} ms, energy {pe['kernel_ms']:.3f} ms")

    # C1 (ii): CG with ILU(0) (Scaling 1, RCMK: the reference's ordering) on the
    # same system, once, outside the timed steps
    so = LA.Solver(Algorithm=1, Preconditioner=6, Scaling=1, Nb_Iter_Max=5000, Stopping_Test=1e-10,
                   Renumbering_Technique=1)
    xo = LA.Vector(fe.N)
    io = LA.solve(A, b, so, xo)
    cg_ilu0 = {"its": io["its"], "status": io["status"], "rel_res": io["rel_res"],
               "setup_ms": round(io["t_setup_ms"], 1), "solve_ms": round(io["t_solve_ms"], 1),
               "ilu_levels": io["ilu_levels"],
               "ilu_apply_ms": round(LA.solver_prec_apply(so, b, LA.Vector(fe.N), reps=5), 3)}
    cg_ilu0["ilu_apply_gbs"] = round(ilu_apply_bytes(fe.N, fe.nnz) / (cg_ilu0["ilu_apply_ms"] * 1e-3) / 1e9, 1)
    log(f"cg_ilu0: its={io['its']} status={io['status']} rel_res={io['rel_res']:.2e} "
        f"{io['t_solve_ms']:.0f} ms ({io['ilu_levels']} ILU levels)")
    so.close()

    # BCGSTAB on the same assembled system and scaling: products to 1e-10 and
    # time per product.  (Sparskit's TFQMR with Jacobi scaling stagnates on
    # this operator from 24^3 cells up, in the reference's own arithmetic too.)
    others = {}
    for name, alg in (("bcgstab_jacobi", 5),):
        so = LA.Solver(Algorithm=alg, Preconditioner=0, Scaling=1, Nb_Iter_Max=20000,
                       Stopping_Test=1e-10, Renumbering_Technique=1)
        xo = LA.Vector(fe.N)
        LA.solve(A, b, so, xo)  # set-up (renumbering, scaling) and graph capture
        io = LA.solve(A, b, so, xo)
        others[name] = {"its": io["its"], "status": io["status"], "rel_res": io["rel_res"],
                        "solve_ms": round(io["t_solve_ms"], 2),
                        "us_per_product": round(1e3 * io["t_solve_ms"] / max(1, io["its"]), 1)}
        log(f"{name}: its={io['its']} status={io['status']} rel_res={io['rel_res']:.2e} "
            f"{io['t_solve_ms']:.1f} ms")

    out = {
        "metric": METRIC,
        "value": round(ms_per_step / 1e3, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"elasticity3d_{args.n}: unit cube {args.n}^3 x 6 P1 tets ({E} elements, "
                         f"{fe.N} DOFs, nnz {fe.nnz}), lognormal E sigma 1, nu 0.3, 4-pt Gauss; "
                         "assemble + CG/Jacobi (Scaling 1) to 1e-10"),
            "elements": E, "dofs": fe.N, "nnz": fe.nnz,
            "solver": "Sparskit CG, Scaling 1, RCMK, FP64",
        },
        "cg_its": int(np.median([i["its"] for i in infos])),
        "rel_res": max(i["rel_res"] for i in infos),
        "solve_ms": round(float(np.mean([i["t_solve_ms"] for i in infos])), 2),
        "setup_ms": round(float(np.mean([i["t_setup_ms"] for i in infos])), 2),
        "assembly_ms": round(asm_s * 1e3, 3),
        "assembly_elems_per_s": round(E / asm_s, 1),
        "assembly_gbs": round(assembly_bytes(pr, fe.N, fe.nnz) / asm_s / 1e9, 1),
        # compulsory bytes (connectivity, coordinates, values and right-hand
        # side once) against the HBM peak, and the counter-measured traffic
        "assembly_roofline": {
            "bound": "hbm", "achieved": round(assembly_bytes(pr, fe.N, fe.nnz) / asm_s / 1e9, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(assembly_bytes(pr, fe.N, fe.nnz) / asm_s / 1e9 / HBM_PEAK_GBS, 4),
            "compulsory": int(assembly_bytes(pr, fe.N, fe.nnz)),
            "traffic": PMC_ASSEMBLY_TRAFFIC.get(args.n),
            "traffic_gbs": (round(PMC_ASSEMBLY_TRAFFIC[args.n] / asm_s / 1e9, 1)
                            if args.n in PMC_ASSEMBLY_TRAFFIC else None),
        },
        "spmv_hbm_frac": roofline["frac"],
        "spmv_natural_order_ms": round(spmv_natural_ms, 5),
        "roofline": roofline,
        "c1_cg_ilu0": cg_ilu0,
        "post": post,
        "other_krylov": others,
        "cpu_baseline": None,
    }
    if not args.no_learned:
        out["learned"] = learned_section()
        out["learned_3d"] = learned3d_section()
    if not args.no_batch:
        out["batch"] = batch_section()
    if not args.no_c2:
        out["c2_mixed_precision"] = c2_section()
    dpr = None
    if not args.no_default_solver:
        out["default_solver"], dpr = default_solver_section()
    if not args.no_cpu:
        try:
            pair = cpu_pair(lambda: cpu_baseline(pr, out["cg_its"], args.cg_sample))
            out["cpu_baseline"] = pair[-1]
            out["cpu_baselines"] = pair
            if dpr is not None:
                out["default_solver"]["cpu"] = default_solver_cpu(dpr)
            if "learned" in out:
                from getdp_amd import learned as L
                out["learned"]["cpu_plain_rcmk"] = cpu_gmres(L.random_field_problem(256, 90000))
            if "learned_3d" in out:
                from getdp_amd import learned3d as L3
                out["learned_3d"]["cpu_plain_rcmk"] = cpu_gmres(
                    L3.random_field_problem(64, 90002), its_gpu=out["learned_3d"]["plain_rcmk"]["its"], sample=40)
            if "c2_mixed_precision" in out:
                from getdp_amd import problems as P
                c2 = out["c2_mixed_precision"]
                cpr = P.curlcurl3d(48)
                c2["cpu_gmres_ilu0"] = cpu_gmres(cpr, its_gpu=c2["fp64"]["its"], sample=40)
                c2["cpu_cg_jacobi"] = cpu_gmres(cpr, its_gpu=c2["fp64_cg_jacobi"]["its"], sample=200,
                                                algorithm=1, precond=0)
            if "batch" in out:
                out["batch"]["cpu"] = batch_cpu()
        except Exception as e:  # the oracle is a checker; report, do not fail
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
