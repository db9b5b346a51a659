"""Headline benchmark of the GPU solve phase on one MI355X.

Workload: the mesh-smoothing system A = I + 10 (D - W) on icosphere(7)
(n = 163,842, nnz(A) = 1,146,882 after duplicate summation), 3 right-hand
sides = vertex positions + N(0, 0.01) noise (seed 0), factored once on the
host (nested dissection + supernodal multifrontal LL').

Paths timed in the same process, inputs resident in HBM:
  refined  FP32 block solve + FP64 refinement steps until the relative
           residual is <= 1e-12 (the steps needed are reported; each step
           reads the residual norms back to the host)
  f64      FP64 block solve
  f32      FP32 block solve alone (reported, below the accuracy bar)
  plain    FP64 row-per-wavefront level-scheduled solve (the reference's
           algorithm on gfx950; the "vs plain" denominator)
  plain32  the same in FP32
The headline is the fastest path whose relative error against the FP64
host solve is <= 1e-10. A step is one solve of the 3 right-hand sides;
`value` = 3 * steps / time. Every path is also timed cold: a 512 MiB
buffer is overwritten before each solve so the factor is not served from
the 256 MB Infinity Cache.

roofline: the dominant pass of the headline path (the triangular pass that
takes longer). A pass at k <= 4 is one launch of the block kernel, whose
leaf phase runs the leaf blocks' tasks (blocks without child blocks, one
wavefront per task) before (forward) or after (backward) the other blocks'
tasks; with the leaf phase off it is two launches. The pass is timed with
HIP events recorded between the two passes of complete solves on the stream
they run on (the null stream); achieved = algorithmic bytes of that pass /
its average time. traffic = the HBM bytes rocprofv3 counted for the pass's
kernels on this workload (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by
the counts of a copy of known size, summed over the pass's kernels), read
from the summary kept beside this file in
profiles/ when it names the same kernels.
cpu_baseline: the oracle's FP64 simplicial solve (oracle/oracle.c,
oracle_host_solve: what cholmod_solve does for a simplicial LL' factor) on
the same factor, one thread, bounded to ~10 s of work; next to it the
all-cores variant, the right-hand-side columns shared out over
min(k, nproc) threads (each column is an independent solve).

--profile runs only the headline path for `steps` solves and then a plain
256 MiB device copy (the known-size kernel that scales the counters); it is
the program the rocprofv3 counter passes run.

After the headline (outside its timed region) the default run starts the
other workloads (--workload precision, batch, grad, reference) as child
processes and records their lines in brief under config.other_workloads;
--no-extra skips them.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "RHS-cols solved/s at ≤1e-10 rel.err vs FP64 CHOLMOD; % HBM roofline; vs plain"
HBM_PEAK_GBS = 8000.0
REFINE_TOL = 1e-12     # relative residual the refined path iterates to
REFINE_MAX_STEPS = 6
TOL = 1e-10


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pass_bytes(s: int, n: int, nnz_l: int, k: int) -> int:
    """Algorithmic bytes of one triangular pass: the factor streamed once
    (value + column index per entry, row pointers), the permutation, and the
    right-hand side read once / result written once (half of B_solve)."""
    return (s + 4) * nnz_l + 4 * (n + 1) + 4 * n + 2 * s * n * k


def time_events(fn, reps: int, warm: int = 3) -> float:
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def measured_traffic(kernels, workload: str, k: int):
    """HBM bytes of one pass (the sum over its kernels of the bytes per
    launch) on `workload` from the committed rocprofv3 counter summary
    (profiles/traffic.json), or None when the summary does not cover every
    kernel of the pass on this workload."""
    path = ROOT / "profiles" / "traffic.json"
    try:
        entries = json.loads(path.read_text())
    except (OSError, ValueError):
        return None
    total = 0
    for kernel in kernels:
        hit = [e for e in entries
               if e.get("kernel") == kernel and e.get("workload") == workload and e.get("k") == k]
        if not hit:
            return None
        total += hit[0].get("bytes_per_launch", 0)
    return total


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--level", type=int, default=7, help="icosphere subdivision level")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--workload", choices=["smoothing", "batch", "precision", "grad", "reference"],
                    default="smoothing")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other workloads (precision, batch, grad, reference) after the headline")
    ap.add_argument("--profile", action="store_true",
                    help="headline path only, then a 256 MiB copy (for rocprofv3 counter passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the solve phase is single-device); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "precision":
        return bench_precision(args)
    if args.workload == "grad":
        return bench_grad(args)
    if args.workload == "reference":
        return bench_reference(args)

    import torch

    import cholespy_amd
    from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, workloads

    dev = torch.device("cuda:0")
    n, values, idx, rhs = workloads.smoothing_system(args.level, 10.0, seed=0)
    k = rhs.shape[1]
    t0 = time.time()
    ii = torch.tensor(idx[0], device=dev)
    jj = torch.tensor(idx[1], device=dev)
    xx = torch.tensor(values, device=dev)
    sf = CholeskySolverF(n, ii, jj, xx, MatrixType.COO)
    sd = CholeskySolverD(n, ii, jj, xx, MatrixType.COO)
    log(f"setup (2 factorisations + analyses + plans): {time.time() - t0:.2f}s")
    # Device autotuning for this system and k (fresh on this box, not
    # cached). The --profile run keeps the defaults, so the profiler sees
    # only the configuration it measures.
    tuned = {}
    if not args.profile:
        t0 = time.time()
        tuned = {"f64": sd.autotune(k, force=True, write=False), "f32": sf.autotune(k, force=True, write=False)}
        log(f"autotune ({time.time() - t0:.1f}s): " + "; ".join(
            f"{p}: {r['ms']:.4f} ms (defaults {r['default_ms']:.4f})" for p, r in tuned.items()))
    fac = sd._factor()
    nnz_l = len(fac["val"])
    nnz_a = len(sd._matrix()["val"]) * 2 - n
    plan = sf._plan_info()

    b64 = torch.tensor(rhs, device=dev, dtype=torch.float64)
    b32 = b64.to(torch.float32)
    x64 = torch.zeros_like(b64)
    x32 = torch.zeros_like(b32)

    # Accuracy reference: FP64 host solve on the same factor.
    from oracle import oracle
    x_ref = oracle.host_solve(fac, rhs)

    def rel_err(x):
        x = x.detach().cpu().numpy().astype(np.float64)
        return float(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref))

    P = {"rows": cholespy_amd._core._PATH_ROWS, "blocks": cholespy_amd._core._PATH_BLOCKS}

    refine_steps = {}

    def run_refined():
        steps, res = sf._solve_refined_tol(b64, x64, REFINE_TOL, REFINE_MAX_STEPS)
        refine_steps["steps"], refine_steps["residual"] = steps, res

    def run_f64():
        sd.solve(b64, x64)

    def run_f32():
        sf.solve(b32, x32)

    # Cold runs: overwrite a 512 MiB buffer (twice the Infinity Cache) before
    # each solve; only the solve is inside the events.
    flush = torch.empty(128 * 1024 * 1024, device=dev)

    def time_cold(fn, reps):
        ts = []
        for _ in range(reps):
            flush.fill_(1.0)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))

    paths = {"refined": (sf, "blocks", run_refined, x64),
             "f64": (sd, "blocks", run_f64, x64),
             "f32": (sf, "blocks", run_f32, x32),
             "plain": (sd, "rows", run_f64, x64),
             "plain32": (sf, "rows", run_f32, x32)}
    if args.profile:
        paths = {"f64": paths["f64"], "refined": paths["refined"]}
    results = {}
    for name, (solver, path, fn, out) in paths.items():
        solver._set_path(P[path])
        fn()
        torch.cuda.synchronize()
        solver._check()
        err = rel_err(out)
        slow = path == "rows"
        reps = 5 if (slow or args.profile) else max(10, args.steps // 4)
        ms = time_events(fn, reps, warm=1 if slow else 3)
        cold = None if args.profile else time_cold(fn, 3 if slow else 20)
        solver._set_path(P["blocks"])
        results[name] = {"ms": ms, "ms_cold": cold, "rel_err": err}
        log(f"{name:8s} {ms:8.4f} ms/solve (cold {cold}) rel.err {err:.2e}")

    ok = [p for p in ("f64", "refined") if results[p]["rel_err"] <= TOL]
    if not ok:
        log("no path reached the accuracy bar")
        ok = ["f64"]
    headline = min(ok, key=lambda p: results[p]["ms"])
    solver, _, fn, xout = paths[headline]

    # Timed region: exactly `steps` solves bracketed by synchronisations.
    for _ in range(args.warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    solver._check()
    ms_step = e0.elapsed_time(e1) / args.steps
    final_err = rel_err(xout)
    value = k * 1000.0 / ms_step
    log(f"timed {args.steps} steps ({headline}): {ms_step:.4f} ms/step "
        f"(wall {wall * 1000 / args.steps:.4f}), rel.err {final_err:.2e}")

    if args.profile:
        # Known-size kernel for scaling FETCH_SIZE / WRITE_SIZE: 256 MiB read,
        # 256 MiB written.
        src = torch.ones(64 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"profile": headline, "ms_per_step": ms_step, "rel_err": final_err}))
        return

    # Roofline of the dominant kernel: each block pass of the headline path.
    # The refined path's passes are FP32 block passes.
    s_bytes = 4 if headline == "refined" else 8
    pb = pass_bytes(s_bytes, n, nnz_l, k)
    solve_bytes = 2 * pb   # B_solve: both triangles, b read, y written+read, x written
    rhs_pass = b32 if headline == "refined" else b64
    out_pass = x32 if headline == "refined" else x64
    # Each pass timed inside complete solves (forward then backward, events
    # between them), so every dependency wait of the real schedule is paid.
    reps = 50
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(reps)]
    for _ in range(3):
        solver._solve_pass(0, rhs_pass, out_pass)
        solver._solve_pass(1, rhs_pass, out_pass)
    torch.cuda.synchronize()
    for e in ev:
        e[0].record()
        solver._solve_pass(0, rhs_pass, out_pass)
        e[1].record()
        solver._solve_pass(1, rhs_pass, out_pass)
        e[2].record()
    torch.cuda.synchronize()
    pass_ms = [float(np.mean([e[0].elapsed_time(e[1]) for e in ev])),
               float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))]
    dom = int(np.argmax(pass_ms))
    achieved = pb / (pass_ms[dom] * 1e-3) / 1e9
    log(f"passes: forward {pass_ms[0] * 1000:.1f} us, backward {pass_ms[1] * 1000:.1f} us, "
        f"algorithmic bytes per pass {pb / 1e6:.1f} MB")
    torch.cuda.synchronize()
    solver._check()
    tname = "<float,4>" if s_bytes == 4 else "<double,4>"
    # A pass is one launch when its leaf tasks run in the block launch's leaf
    # phase (the default), else a leaf launch and a block launch.
    fused = tuned.get("f32" if s_bytes == 4 else "f64", {}).get("fused", 3)
    if (fused >> (0 if dom == 0 else 1)) & 1:
        pass_kernels = ["block_forward" + tname] if dom == 0 else ["block_backward" + tname]
    else:
        pass_kernels = ["leaf_forward" + tname, "block_forward" + tname] if dom == 0 else \
                       ["block_backward" + tname, "leaf_backward" + tname]
    kernel = ("forward" if dom == 0 else "backward") + " pass: " + " + ".join(pass_kernels)
    workload_name = f"mesh smoothing I+10(D-W) on icosphere({args.level})"
    traffic = measured_traffic(pass_kernels, workload_name, k)

    # The reference's own published benchmark (tests/Benchmark.ipynb:238-275,
    # README.md:24-33): icosphere(6), A = I + 20 (D - W) (the last lambda of
    # the timed loop), 128 right-hand sides U[0,1), CholeskySolverF.solve;
    # cholespy on a TITAN RTX: 6.42 ms per solve.
    ref_bench = None
    try:
        n6, f6 = workloads.icosphere(6)
        v6, i6 = workloads.mesh_laplacian(n6, f6, 20.0)
        s6 = CholeskySolverF(n6, torch.tensor(i6[0], device=dev), torch.tensor(i6[1], device=dev),
                             torch.tensor(v6, device=dev), MatrixType.COO)
        b6 = torch.rand(n6, 128, device=dev)
        x6 = torch.zeros_like(b6)
        ms6_default = time_events(lambda: s6.solve(b6, x6), 20)
        tune6 = s6.autotune(128, force=True, write=False)
        ms6 = time_events(lambda: s6.solve(b6, x6), 20)
        ms6z = time_events(lambda: (torch.zeros_like(b6), s6.solve(b6, x6)), 20)
        s6._check()
        # The published "analysis" bar is the constructor (CHOLMOD factor +
        # level analysis): 331.67 ms. The timed loop builds a new solver for
        # every lambda; refactorize() replaces that for a fixed pattern.
        ti6 = [torch.tensor(i6[0], device=dev), torch.tensor(i6[1], device=dev)]
        construct = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s6b = CholeskySolverF(n6, ti6[0], ti6[1], torch.tensor(v6, device=dev), MatrixType.COO)
            torch.cuda.synchronize()
            construct.append((time.perf_counter() - t0) * 1e3)
            del s6b
        refac = []
        for lam in (18.0, 19.0, 20.0):
            vl = torch.tensor(workloads.mesh_laplacian(n6, f6, lam)[0], device=dev)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s6.refactorize(vl)
            torch.cuda.synchronize()
            refac.append((time.perf_counter() - t0) * 1e3)
        ref_bench = {"workload": "icosphere(6), n=40962, lambda=20, k=128, float32", "ms_solve": ms6,
                     "ms_solve_defaults": ms6_default,
                     "autotune": {kk: v for kk, v in tune6.items() if kk != "key"},
                     "ms_solve_with_zeros_like": ms6z, "reference_ms_titan_rtx": 6.42,
                     "speedup_vs_reference": 6.42 / ms6z,
                     "ms_construct": float(np.median(construct)), "ms_refactorize": float(np.median(refac)),
                     "reference_ms_construct": 331.67}
        log(f"reference benchmark config: {ms6:.3f} ms/solve ({ms6z:.3f} with zeros_like) vs 6.42 ms; "
            f"construction {np.median(construct):.1f} ms, refactorize {np.median(refac):.1f} ms vs 331.67 ms")
        del s6
    except Exception as e:  # pragma: no cover - reported, not fatal
        log(f"reference benchmark config failed: {e}")

    # CPU baseline: oracle FP64 simplicial solve, one thread, ~cpu_seconds of work.
    cpu = None
    if not args.no_cpu:
        times = []
        t_end = time.perf_counter() + args.cpu_seconds
        for _ in range(3):
            oracle.host_solve(fac, rhs)
        while time.perf_counter() < t_end or len(times) < 5:
            c0 = time.perf_counter()
            oracle.host_solve(fac, rhs)
            times.append(time.perf_counter() - c0)
        med = float(np.median(times))
        nproc = os.cpu_count() or 1
        threads = max(1, min(k, nproc))
        ptimes = []
        t_end = time.perf_counter() + args.cpu_seconds / 2
        oracle.host_solve_parallel(fac, rhs, threads)
        while time.perf_counter() < t_end or len(ptimes) < 5:
            c0 = time.perf_counter()
            oracle.host_solve_parallel(fac, rhs, threads)
            ptimes.append(time.perf_counter() - c0)
        pmed = float(np.median(ptimes))
        cpu = {"value": k / med, "unit": "RHS-cols/s", "cores": 1, "kind": "port",
               "sample": f"{len(times)} FP64 solves of the same system (k={k}) on the same factor, "
                         f"median {med * 1000:.2f} ms, single thread",
               "all_cores": {"value": k / pmed, "unit": "RHS-cols/s", "cores": threads, "nproc": nproc,
                             "sample": f"{len(ptimes)} solves, the {k} columns over {threads} threads "
                                       f"(nproc {nproc}), median {pmed * 1000:.2f} ms"}}
        log(f"cpu baseline: {med * 1000:.2f} ms/solve ({len(times)} runs); "
            f"{threads} threads: {pmed * 1000:.2f} ms (nproc {nproc})")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "RHS-cols/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if headline == "refined" else "f64",
        "data": "synthetic",
        "config": {
            "workload": workload_name,
            "n": n, "nnz_A": nnz_a, "nnz_L": nnz_l, "k": k,
            "path": "blocks FP32 + FP64 refinement" if headline == "refined" else "blocks FP64",
            "refine_tol": REFINE_TOL,
            "refine_steps": refine_steps.get("steps"),
            "refine_residual": refine_steps.get("residual"),
            "rel_err": final_err,
            "solves_per_s": 1000.0 / ms_step,
            "ms": {name: r["ms"] for name, r in results.items()},
            "ms_cold": {name: r["ms_cold"] for name, r in results.items()},
            "rel_errs": {name: r["rel_err"] for name, r in results.items()},
            "vs_plain": results["plain"]["ms"] / ms_step,
            "refined_vs_plain": results["plain"]["ms"] / results["refined"]["ms"],
            "refined_vs_f64": results["f64"]["ms"] / results["refined"]["ms"],
            "solve_roofline_frac": solve_bytes / (ms_step * 1e-3) / (HBM_PEAK_GBS * 1e9),
            "block_plan": plan,
            "autotune": {p: {kk: v for kk, v in r.items() if kk != "key"} for p, r in tuned.items()},
            "reference_benchmark": ref_bench,
            "other_workloads": other_workloads(args),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "pass_us": [pass_ms[0] * 1000, pass_ms[1] * 1000],
            "kernels": pass_kernels,
            "bytes_per_launch": pb,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def other_workloads(args):
    """The other configurations, each in a child process after the headline
    (outside its timed region): their bench lines in brief, so one default
    run records every configuration's numbers."""
    if args.no_extra or args.profile:
        return None
    import subprocess

    runs = {"precision": ["--steps", "100"], "batch": ["--steps", "50"], "grad": ["--steps", "50"],
            "reference": ["--steps", "200"]}
    keep = {"precision": ("refine_steps", "ms", "refined_vs_f64", "adaptive"),
            "batch": ("n_total", "sum_k", "ms_batched", "ms_graph", "ms_per_system", "batched_vs_per_system",
                      "rel_err_f32_vs_oracle"),
            "grad": ("ms_f32", "ms_f64", "entry_grad_rel_diff_f32_vs_f64"),
            "reference": ("ms_defaults", "rel_err_vs_fp64_host", "reference_ms_titan_rtx")}
    out = {}
    for name, extra in runs.items():
        t0 = time.time()
        try:
            r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--workload", name, *extra],
                               capture_output=True, text=True, timeout=600)
            line = [l for l in r.stdout.strip().splitlines() if l.startswith("{")][-1] if r.returncode == 0 else None
        except (subprocess.TimeoutExpired, IndexError) as e:
            r, line = None, None
            log(f"{name} workload: {e}")
        if line is None:
            out[name] = {"error": (r.stderr[-400:] if r is not None else "no result")}
            continue
        d = json.loads(line)
        cfg = d.get("config", {})
        out[name] = {"workload": cfg.get("workload"), "value": d.get("value"), "unit": d.get("unit"),
                     "ms_per_step": d.get("ms_per_step"), "dtype": d.get("dtype"),
                     "vs_baseline": d.get("vs_baseline"), "roofline": d.get("roofline"),
                     **{kk: cfg[kk] for kk in keep[name] if kk in cfg},
                     "wall_s": round(time.time() - t0, 1)}
        log(f"{name} workload: {d.get('ms_per_step'):.4f} ms per step ({time.time() - t0:.0f} s)")
    return out


def bench_precision(args):
    """3D 7-point graph Laplacian (Neumann) on a 64^3 grid + 1e-4 I (n =
    262,144), factored once in FP64 on the host and solved in FP32 on the GPU
    for 32 right-hand sides ~ N(0, 1) (seed 0), with FP64 residual
    refinement until the relative residual max_j ||b_j - A x_j|| / ||b_j||
    is <= 1e-12. Reports the refinement steps needed, the error against the
    FP64 host solve on the same factor, and the time per refined solve next
    to the FP32 and FP64 block solves."""
    import scipy.sparse as sp
    import torch

    from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, workloads
    from oracle import oracle

    dev = torch.device("cuda:0")
    m = 64
    n = m ** 3
    k = 32
    values, idx = workloads.grid3d_laplacian(m, 1e-4)
    rhs = np.random.default_rng(0).standard_normal((n, k))
    t0 = time.time()
    ii, jj, xx = (torch.tensor(idx[0], device=dev), torch.tensor(idx[1], device=dev),
                  torch.tensor(values, device=dev))
    sf = CholeskySolverF(n, ii, jj, xx, MatrixType.COO)
    log(f"FP32 solver (FP64 host factor + plan): {time.time() - t0:.1f}s, plan {sf._plan_info()}")
    if args.profile:
        # The counter passes: `steps` FP32 panel solves with the default
        # schedule, then the 256 MiB copy that scales the counters.
        b32 = torch.tensor(rhs, device=dev, dtype=torch.float32)
        x32 = torch.zeros_like(b32)
        ms = time_events(lambda: sf.solve(b32, x32), max(1, args.steps))
        sf._check()
        src = torch.ones(64 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"profile": "f32 panel solve", "ms_per_step": ms}))
        return
    t0 = time.time()
    tune = sf.autotune(k, force=True, write=False)
    log(f"autotune ({time.time() - t0:.1f}s): {tune['ms']:.4f} ms per FP32 solve (defaults {tune['default_ms']:.4f})")
    fac = sf._factor()
    A = sp.coo_matrix((values, (idx[0], idx[1])), shape=(n, n)).tocsr()
    t0 = time.time()
    x_ref = oracle.host_solve(fac, rhs)
    log(f"FP64 host solve (oracle): {time.time() - t0:.1f}s")
    bnorm = np.linalg.norm(rhs, axis=0)

    def residual(x):
        return float((np.linalg.norm(rhs - A @ x, axis=0) / bnorm).max())

    def rel_err(x):
        return float(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref))

    b64 = torch.tensor(rhs, device=dev)
    x64 = torch.zeros_like(b64)
    b32 = b64.float()
    x32 = torch.zeros_like(b32)
    sf.solve(b32, x32)
    torch.cuda.synchronize()
    sf._check()
    x32h = x32.cpu().numpy().astype(np.float64)
    steps = None
    trail = [{"steps": 0, "residual": residual(x32h), "rel_err": rel_err(x32h)}]
    for it in range(1, 7):
        sf._solve_refined(b64, x64, it)
        torch.cuda.synchronize()
        sf._check()
        xh = x64.cpu().numpy()
        trail.append({"steps": it, "residual": residual(xh), "rel_err": rel_err(xh)})
        log(f"refinement steps {it}: residual {trail[-1]['residual']:.2e} rel.err {trail[-1]['rel_err']:.2e}")
        if trail[-1]["residual"] <= 1e-12:
            steps = it
            break
    reps = max(5, args.steps // 20)
    # The timed path: refinement to the same bar, the stopping test on the
    # device (the residual norms read back before each step).
    tol_steps, tol_res = sf._solve_refined_tol(b64, x64, 1e-12, 6)
    torch.cuda.synchronize()
    xt = x64.cpu().numpy()
    log(f"adaptive refinement: {tol_steps} steps, device residual {tol_res:.2e}, host residual {residual(xt):.2e}")
    ms_refined = time_events(lambda: sf._solve_refined_tol(b64, x64, 1e-12, 6), reps)
    ms_f32 = time_events(lambda: sf.solve(b32, x32), reps)
    sf._check()
    del sf
    torch.cuda.empty_cache()
    t0 = time.time()
    sd = CholeskySolverD(n, ii, jj, xx, MatrixType.COO)
    log(f"FP64 solver: {time.time() - t0:.1f}s")
    ms_f64 = time_events(lambda: sd.solve(b64, x64), reps)
    sd._check()
    x64h = x64.cpu().numpy()
    out = {
        "metric": METRIC,
        "value": k * 1000.0 / ms_refined,
        "unit": "RHS-cols/s",
        "n_gpus": 1, "steps": reps, "warmup": 3, "ms_per_step": ms_refined,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "3D 7-point Neumann Laplacian on 64^3 + 1e-4 I, 32 RHS N(0,1)",
            "n": n, "k": k, "nnz_L": int(len(fac["val"])),
            "kappa_estimate": (12.0 + 1e-4) / 1e-4,
            "refine_steps": steps, "refinement": trail,
            "adaptive": {"steps": tol_steps, "device_residual": tol_res, "residual": residual(xt),
                         "rel_err": rel_err(xt)},
            "ms": {"refined": ms_refined, "f32": ms_f32, "f64": ms_f64},
            "f64": {"residual": residual(x64h), "rel_err": rel_err(x64h)},
            "refined_vs_f64": ms_f64 / ms_refined,
            "autotune_f32": {kk: v for kk, v in tune.items() if kk != "key"},
        },
    }
    print(json.dumps(out))


def bench_grad(args):
    """Differentiable solve on icosphere(6) (n = 40,962), A = I + 10 (D - W),
    3 right-hand sides, FP32 solver: loss = ||x - t||^2 and its gradients
    with respect to b and every stored entry of A (adjoint solve + HIP
    scatter). A step is one forward solve + backward; the FP32 gradients
    are compared with the FP64 solver's."""
    import torch

    from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, workloads
    from cholespy_amd.autograd import coo_indices, solve

    dev = torch.device("cuda:0")
    n, faces = workloads.icosphere(6)
    values, idx = workloads.mesh_laplacian(n, faces, 10.0)
    rng = np.random.default_rng(0)
    b = rng.standard_normal((n, 3))
    t = rng.standard_normal((n, 3))
    ii, jj = torch.tensor(idx[0], device=dev), torch.tensor(idx[1], device=dev)
    rows, cols = coo_indices(n, ii, jj, MatrixType.COO)
    out = {}
    for name, cls, dt in (("f32", CholeskySolverF, torch.float32), ("f64", CholeskySolverD, torch.float64)):
        vals = torch.tensor(values, device=dev, requires_grad=True)
        solver = cls(n, ii, jj, vals.detach(), MatrixType.COO)
        bt = torch.tensor(b, device=dev, dtype=dt, requires_grad=True)
        tt = torch.tensor(t, device=dev, dtype=dt)

        def step():
            vals.grad = None
            bt.grad = None
            x = solve(solver, bt, vals, rows, cols)
            ((x - tt) ** 2).sum().backward()

        ms = time_events(step, max(20, args.steps // 4))
        solver._check()
        out[name] = {"ms": ms, "gv": vals.grad.double().cpu().numpy(), "gb": bt.grad.double().cpu().numpy()}
        log(f"{name}: {ms:.3f} ms per forward + backward")
    rel = float(np.abs(out["f32"]["gv"] - out["f64"]["gv"]).max() / np.abs(out["f64"]["gv"]).max())
    ms = out["f32"]["ms"]
    print(json.dumps({
        "metric": METRIC, "value": 3 * 1000.0 / ms, "unit": "RHS-cols/s (forward + backward)",
        "n_gpus": 1, "steps": max(20, args.steps // 4), "warmup": 3, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "differentiable solve, icosphere(6), lambda=10, k=3, loss ||x - t||^2",
                   "n": n, "ms_f32": ms, "ms_f64": out["f64"]["ms"],
                   "entry_grad_rel_diff_f32_vs_f64": rel},
    }))


def bench_reference(args):
    """The reference's own published benchmark configuration alone
    (tests/Benchmark.ipynb:238-275, README.md:24-33): icosphere(6) (n =
    40,962), A = I + 20 (D - W), 128 right-hand sides U[0, 1) (seed 0),
    CholeskySolverF.solve; cholespy on a TITAN RTX: 6.42 ms per solve. The
    solver is autotuned first; with --profile, `steps` solves with the
    default schedule, then a 256 MiB copy (for rocprofv3 counter passes)."""
    import torch

    from cholespy_amd import CholeskySolverF, MatrixType, workloads
    from oracle import oracle

    dev = torch.device("cuda:0")
    n, faces = workloads.icosphere(6)
    values, idx = workloads.mesh_laplacian(n, faces, 20.0)
    k = 128
    s = CholeskySolverF(n, torch.tensor(idx[0], device=dev), torch.tensor(idx[1], device=dev),
                        torch.tensor(values, device=dev), MatrixType.COO)
    g = torch.Generator().manual_seed(0)
    b = torch.rand(n, k, generator=g).to(dev)
    x = torch.zeros_like(b)
    if args.profile:
        ms = time_events(lambda: s.solve(b, x), max(1, args.steps))
        s._check()
        src = torch.ones(64 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"profile": "icosphere(6) k=128 FP32 panel solve", "ms_per_step": ms}))
        return
    ms_default = time_events(lambda: s.solve(b, x), max(5, args.steps // 4))
    tune = s.autotune(k, force=True, write=False)
    ms = time_events(lambda: s.solve(b, x), args.steps)
    s._check()
    ref = oracle.host_solve(s._factor(), b.cpu().double().numpy())
    err = float(np.abs(x.cpu().double().numpy() - ref).max() / np.abs(ref).max())
    entries = int(s._plan_info()["m_entries"])
    flops = 4.0 * entries * k          # both passes: 2 flops per M entry and column
    tflops = flops / (ms * 1e-3) / 1e12
    log(f"icosphere(6) k=128 FP32: {ms:.4f} ms per solve (defaults {ms_default:.4f}), rel.err {err:.1e}, "
        f"{tflops:.1f} TFLOP/s, {6.42 / ms:.1f}x the published 6.42 ms")
    print(json.dumps({
        "metric": "ms per solve, cholespy's published benchmark configuration", "value": ms, "unit": "ms",
        "n_gpus": 1, "steps": args.steps, "warmup": 3, "ms_per_step": ms,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": 6.42 / ms, "dtype": "f32",
        "data": "synthetic",
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": 157.3, "unit": "TFLOP/s",
                     "frac": tflops / 157.3, "traffic": None},
        "config": {"workload": "icosphere(6), n=40962, lambda=20, k=128, float32, b ~ U[0,1)",
                   "m_entries": entries, "ms_defaults": ms_default, "rel_err_vs_fp64_host": err,
                   "reference_ms_titan_rtx": 6.42,
                   "autotune": {kk: v for kk, v in tune.items() if kk != "key"}},
    }))


def bench_batch(args):
    """256 independent grid systems (n log-uniform in [1k, 50k], k_i uniform
    in {1..128}), FP32: one launch per pass for the whole batch, the same
    replayed from a captured hipGraph, and per-system launches (one solve call
    per system on its own solver). value = sum(k_i) * steps / time."""
    import torch

    from cholespy_amd import BatchedSolverF, CholeskySolverF, MatrixType, workloads
    from oracle import oracle

    t0 = time.time()
    systems = workloads.batch_systems(count=args.batch_count, seed=0)
    ks = [k for *_, k in systems]
    batch = BatchedSolverF([(n, idx[0], idx[1], v, MatrixType.COO) for n, v, idx, _ in systems])
    log(f"batch setup: {len(systems)} systems, n total {sum(batch.sizes)}, "
        f"sum k {sum(ks)}, {time.time() - t0:.1f}s; plan {batch._plan_info()}")
    rng = np.random.default_rng(1)
    bs = [rng.uniform(-1, 1, (n, k)).astype(np.float32) for (n, *_), k in zip(systems, ks)]
    b = torch.tensor(np.concatenate([x.ravel() for x in bs]), device="cuda")
    x = torch.zeros_like(b)
    run = lambda: batch.solve(ks, b, x)
    run()
    torch.cuda.synchronize()
    batch._check()
    # accuracy: a few systems against the oracle on their factors
    xs = x.cpu().numpy()
    offs = np.concatenate([[0], np.cumsum([n * k for (n, *_), k in zip(systems, ks)])])
    worst = 0.0
    for i in range(0, len(systems), max(1, len(systems) // 8)):
        n = systems[i][0]
        ref = oracle.solve(batch._factor(i), bs[i], np.float32)
        got = xs[offs[i]:offs[i + 1]].reshape(n, ks[i])
        worst = max(worst, float(np.abs(got - ref).max() / np.abs(ref).max()))
    ms_batch = time_events(run, args.steps)
    # hipGraph: capture one batched solve on torch's capture stream, replay.
    ms_graph = None
    try:
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            batch.solve(ks, b, x, torch.cuda.current_stream().cuda_stream)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            batch.solve(ks, b, x, torch.cuda.current_stream().cuda_stream)
        ms_graph = time_events(g.replay, args.steps)
        batch._check()
    except Exception as e:  # pragma: no cover - reported, not fatal
        log(f"graph capture failed: {e}")
    # per-system launches on separate solvers (same factors: same input)
    singles = []
    for n, v, idx, _ in systems:
        singles.append(CholeskySolverF(n, torch.tensor(idx[0], device="cuda"), torch.tensor(idx[1], device="cuda"),
                                       torch.tensor(v, device="cuda"), MatrixType.COO))
    views_b = [b[offs[i]:offs[i + 1]].view(systems[i][0], ks[i]) for i in range(len(systems))]
    views_x = [x[offs[i]:offs[i + 1]].view(systems[i][0], ks[i]) for i in range(len(systems))]

    def run_singles():
        for sv, bv, xv in zip(singles, views_b, views_x):
            sv.solve(bv, xv)
    ms_singles = time_events(run_singles, max(3, args.steps // 20))
    best = min(t for t in (ms_batch, ms_graph) if t is not None)
    cols = sum(ks)
    log(f"batched {ms_batch:.3f} ms, graph {ms_graph}, per-system {ms_singles:.3f} ms, worst rel err {worst:.2e}")
    out = {
        "metric": METRIC,
        "value": cols * 1000.0 / best,
        "unit": "RHS-cols/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": best,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batch of {len(systems)} grid systems, n in [1k,50k], k_i in 1..128",
                   "n_total": int(sum(batch.sizes)), "sum_k": cols, "ms_batched": ms_batch,
                   "ms_graph": ms_graph, "ms_per_system": ms_singles,
                   "batched_vs_per_system": ms_singles / best, "rel_err_f32_vs_oracle": worst,
                   "plan": batch._plan_info()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
