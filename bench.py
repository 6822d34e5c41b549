"""Benchmark of the per-Newton-step hot path on one MI355X.

Workload: the kernel-isolation PCG shape -- a block-sparse SPD system on a
1024x1024 triangulated vertex grid (1,048,576 rows of 3x3 fp32 blocks, about
7 blocks per row, 4 stored in the upper triangle), each edge block
h^2 * Q^T diag(l) Q with l log-uniform in [1e2, 1e6], plus mass * I; random
unit-normal right-hand side. One step = one block-Jacobi PCG solve to
|r|_1/|r0|_1 < 1e-3 from x0 = 0, the linear solve of a Newton step. Inputs
are resident in HBM before timing starts.

Prints one JSON line. roofline: the SpMV kernel (the dominant kernel of an
iteration), algorithmic bytes n*(40*s + 28) per launch with s stored upper
blocks per row, over its mean launch duration from HIP events on the stream
it runs on. cpu_baseline: the oracle (reference-algorithm CPU restatement,
OpenMP) solving the same system on this host's cores.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ppf-contact-solver_amd"))

METRIC = ("ms/Newton step + PCG iters to 1e-3 rel residual, learned precond "
          "vs Jacobi")
HBM_PEAK_GBS = 8000.0


def traffic_from_profiles():
    """HBM bytes per SpMV launch from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "pcg_spmv_grid1024.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def cpu_baseline(sys_np, tol):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    r = oracle.pcg(sys_np, tol=tol)
    dt = time.perf_counter() - t0
    return {"value": dt * 1e3, "unit": "ms/solve", "cores": threads, "kind": "port",
            "sample": f"one full solve of the same 1,048,576-row system "
                      f"({r['iter']} iterations, status {r['status']}); "
                      "reference-algorithm CPU restatement, not the reference binary"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="first run a 256 MiB device copy (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from ppf_amd import scenes
    from ppf_amd.solver import BlockJacobiPCG

    dev = torch.device("cuda:0")
    if args.copy_calib:
        # A plain copy of known size: 256 MiB read + 256 MiB written. Its
        # FETCH_SIZE / WRITE_SIZE calibrate the counters of the kernels below.
        a = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    s = scenes.grid_spd_system(args.res)
    n, nnz = s["n"], len(s["col_index"])
    t = {k: torch.from_numpy(np.ascontiguousarray(s[k])).to(dev).reshape(-1)
         for k in ("row_offset", "col_index", "value", "cdiag", "b")}
    plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
    x = torch.zeros_like(t["b"])

    def step():
        x.zero_()
        return plan.solve(t["value"], t["cdiag"], t["b"], x, tol=args.tol)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = []
    for _ in range(args.steps):
        res = step()
        iters.append(res.iter)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    assert res.status == 0, f"PCG did not converge: status {res.status}"

    spmv_ms = plan.time_spmv(t["value"], t["cdiag"], t["b"], reps=50)
    s_blocks = nnz / n
    alg_bytes = n * (40.0 * s_blocks + 28.0)
    achieved = alg_bytes / (spmv_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles()

    out = {
        "metric": METRIC,
        "value": round(ms, 4),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "pcg_iters": int(np.median(iters)),
        "precond": "block-jacobi",
        "config": {"workload": f"kernel-isolation PCG: {args.res}x{args.res} "
                               f"triangulated grid, {n} rows of 3x3 fp32 blocks "
                               f"({nnz} stored upper blocks), block-Jacobi to "
                               f"{args.tol:g} L1 relative residual",
                   "rows": n, "stored_blocks": nnz, "tol": args.tol},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "pcg spmv (k_spmv_ell<true,3> on this grid)",
                     "kernel_ms": round(spmv_ms, 5),
                     "alg_bytes_per_launch": alg_bytes},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(s, args.tol)
    plan.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
