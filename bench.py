"""Benchmark of the per-Newton-step hot path on one MI355X.

Default workload (--workload drape-stack): the stacked-layer contact scene --
8 cloth sheets of 256 x 256 vertices (1,040,400 triangles), stacked 2 mm
apart with seeded 1e-4 m vertex noise, falling onto a cylinder of radius
0.2 m spinning at 2 rad/s (its vertices pinned and driven kinematically),
fp32, contact gap 1e-3 m, friction 0.5, dt = 1/60 s. W untimed time steps,
then K timed time steps; every Newton step of those K steps is timed (elastic
and contact assembly, block-Jacobi PCG to 1e-3 L1 relative residual, CCD
line search). value = wall milliseconds per Newton step with the
reference's block-Jacobi preconditioner; the PCG iterations per solve are
reported beside it. "learned": the learned preconditioner, trained first on
Newton systems of the same scene family (8 x 64x64 sheets, other seeds), on
the same time steps and on the same last Newton system. Inputs are resident
in HBM before timing.

--workload pcg-grid: the kernel-isolation PCG system (1024 x 1024 grid, ~1M
rows of 3x3 blocks); one step = one block-Jacobi PCG solve.

roofline: the PCG SpMV (the kernel every Newton step runs most: once per PCG
iteration), on the last Newton system of the timed steps; algorithmic bytes
n * (40 s + 28) per launch (s stored upper blocks per row) over the mean
launch duration from HIP events on the plan's stream. cpu_baseline: the
oracle (the CPU restatement of the reference algorithm, OpenMP) on a bounded
sample of the same workload, on this host's cores.
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


def traffic_from_profiles(name, rows, stored_blocks):
    """HBM bytes per SpMV launch from a committed rocprofv3 PMC summary of the
    same system: the summary's `traffic:` line names the rows and stored
    blocks of the system it measured; any other system gets None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for m in re.finditer(r"traffic: rows=(\d+) stored_blocks=(\d+) bytes_per_launch=([0-9.eE+]+)",
                         open(path).read()):
        if int(m.group(1)) == rows and int(m.group(2)) == stored_blocks:
            return float(m.group(3))
    return None


def host_cpu():
    """nproc, CPU model, sockets, cores per socket, threads per core of this host."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        text = open("/proc/cpuinfo").read()
        model = re.search(r"model name\s*:\s*(.*)", text)
        info["model"] = model.group(1).strip() if model else None
        sockets = set(re.findall(r"physical id\s*:\s*(\d+)", text))
        cores = re.search(r"cpu cores\s*:\s*(\d+)", text)
        sib = re.search(r"siblings\s*:\s*(\d+)", text)
        info["sockets"] = len(sockets) or None
        if cores:
            info["cores_per_socket"] = int(cores.group(1))
            if sib:
                info["threads_per_core"] = int(sib.group(1)) // int(cores.group(1))
    except OSError:
        pass
    return info


def progress(msg):
    """One line per unit of work on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def oracle_module():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    return oracle, threads


def spmv_roofline(plan, value, cdiag, b, n, nnz, kernel, traffic):
    spmv_ms = plan.time_spmv(value, cdiag, b, reps=50)
    s_blocks = nnz / n
    alg_bytes = n * (40.0 * s_blocks + 28.0)
    achieved = alg_bytes / (spmv_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": kernel, "kernel_ms": round(spmv_ms, 5),
            "alg_bytes_per_launch": alg_bytes, "rows": n, "stored_blocks": nnz}


def head(args, value, ms, workload, extra):
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "ms/Newton step" if args.workload == "drape-stack" else "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "precond": "block-jacobi",
        "config": workload,
    }
    out.update(extra)
    return out


def run_drape(args, dev):
    import numpy as np
    import torch

    from ppf_amd import scenes
    from ppf_amd.solver import BlockJacobiPCG
    from ppf_amd.stepper import Stepper, scene_to_device

    sc = scenes.stacked_drape_scene(res=args.res, layers=args.layers)
    ns = args.layers * args.res * args.res
    t = scene_to_device(sc, dev)
    t["fix_delta"] = torch.zeros_like(t["mesh"]["x"])
    gpu = Stepper(t)
    p = scenes.step_params()
    clock = [0.0]

    def one_step():
        # the cylinder aimed at the step's end, with its displacement over the
        # step (the kinematic pin's step_delta)
        now = scenes.spin_fix_positions(sc, clock[0])[ns:].copy()
        aim = scenes.spin_fix_positions(sc, clock[0] + p["dt"])[ns:].copy()
        t["fix_pos"].view(-1, 3)[ns:] = torch.from_numpy(aim).to(dev)
        t["fix_delta"].view(-1, 3)[ns:] = torch.from_numpy(aim - now).to(dev)
        r = gpu.advance(p)
        if r["status"] != 0:
            raise RuntimeError(f"step failed: {r}")
        p["prev_dt"] = r["dt"]
        clock[0] += r["dt"]
        progress(f"time step: newton {r['newton_steps']} pcg {r['pcg_iter_total']} pairs "
                 f"{r['pairs_max']} ms {r['ms_total']:.1f}")
        return r

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    # the state the timed window starts from (the CPU baseline steps from it)
    start = {"x": t["mesh"]["x"].cpu().numpy().reshape(-1, 3).copy(),
             "prev": t["prev"].cpu().numpy().reshape(-1, 3).copy(),
             "clock": clock[0], "prev_dt": p["prev_dt"]}
    newton = pcg = pairs = 0
    phase = {"elastic": 0.0, "contact": 0.0, "solve": 0.0, "line_search": 0.0}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = one_step()
        newton += r["newton_steps"]
        pcg += r["pcg_iter_total"]
        pairs = max(pairs, r["pairs_max"])
        for k in phase:
            phase[k] += r["ms_" + k]
    torch.cuda.synchronize()
    total_ms = (time.perf_counter() - t0) * 1e3
    per_newton = total_ms / newton
    sysd = gpu.system()
    learned = None
    if args.precond == "both":
        learned = run_learned(args, dev, sysd)
    n, nnz = sysd["row_offset"].numel() - 1, sysd["col_index"].numel()
    plan = BlockJacobiPCG(sysd["row_offset"], sysd["col_index"])
    # the row order the stepper solves in (the kernel its iterations run)
    plan.set_order(gpu.solve_order())
    plan.repattern(sysd["row_offset"], sysd["col_index"])
    roof = spmv_roofline(plan, sysd["value"], sysd["cdiag"], sysd["b"], n, nnz,
                         "pcg spmv on the last Newton system of the timed steps",
                         traffic_from_profiles("pcg_spmv_drape_stack.md", n, nnz))
    plan.close()
    walks = walk_stats_step(one_step)
    gpu.close()
    workload = {"workload": f"stacked-layer drape: {args.layers} sheets of {args.res}x{args.res} "
                            "vertices 2 mm apart (noise 1e-4 m) over a cylinder r = 0.2 m "
                            "spinning at 2 rad/s, gap 1e-3 m, friction 0.5, dt 1/60 s; time "
                            f"steps {args.warmup + 1}..{args.warmup + args.steps} of the drape",
                "vertices": int(sc["mesh"]["nv"]), "faces": int(sc["mesh"]["nf"]),
                "time_steps": args.steps, "newton_steps": newton,
                "tol": 1e-3}
    extra = {"pcg_iters": round(pcg / newton, 1),
             "learned": learned,
             "phase_ms_per_newton_step": {k: round(v / newton, 3) for k, v in phase.items()},
             "walks": walks,
             "max_contact_pairs": pairs,
             "roofline": roof, "cpu_baseline": None}
    out = head(args, per_newton, per_newton, workload, extra)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_drape(args, start)
    return out


def run_learned(args, dev, last_system):
    """The learned preconditioner: trained on Newton systems of the same
    scene family at 64 x 64 per sheet on other seeds (the block-Jacobi solver
    generates them), then (a) the last Newton system of the timed block-Jacobi
    steps solved both ways, (b) the same time steps re-run with the learned
    factor in every Newton solve."""
    import numpy as np
    import torch

    from ppf_amd import lpc_train, scenes
    from ppf_amd.stepper import Stepper, scene_to_device

    t0 = time.perf_counter()
    systems = lpc_train.collect_drape_systems(dev, res=64, layers=args.layers, seeds=(0, 1, 2),
                                              steps=12)
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    lpc, hist = lpc_train.train(systems, steps=args.lpc_steps, lr=2e-3, log=progress)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    same = lpc_train.compare(lpc, [last_system])[0]
    del systems
    # the timed steps again, learned factor in every Newton solve
    sc = scenes.stacked_drape_scene(res=args.res, layers=args.layers)
    ns = args.layers * args.res * args.res
    t = scene_to_device(sc, dev)
    t["fix_delta"] = torch.zeros_like(t["mesh"]["x"])
    gpu = Stepper(t)
    gpu.set_preconditioner(lpc)
    p = scenes.step_params()
    clock = [0.0]

    def one_step():
        now = scenes.spin_fix_positions(sc, clock[0])[ns:].copy()
        aim = scenes.spin_fix_positions(sc, clock[0] + p["dt"])[ns:].copy()
        t["fix_pos"].view(-1, 3)[ns:] = torch.from_numpy(aim).to(dev)
        t["fix_delta"].view(-1, 3)[ns:] = torch.from_numpy(aim - now).to(dev)
        r = gpu.advance(p)
        if r["status"] != 0:
            raise RuntimeError(f"learned step failed: {r}")
        p["prev_dt"] = r["dt"]
        clock[0] += r["dt"]
        progress(f"learned time step: newton {r['newton_steps']} pcg {r['pcg_iter_total']} "
                 f"ms {r['ms_total']:.1f}")
        return r

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    newton = pcg = fallbacks = 0
    solve_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = one_step()
        newton += r["newton_steps"]
        pcg += r["pcg_iter_total"]
        fallbacks += r["lpc_fallbacks"]
        solve_ms += r["ms_solve"]
    torch.cuda.synchronize()
    total = (time.perf_counter() - t0) * 1e3
    gpu.close()
    lpc.close()
    return {"ms_per_newton_step": round(total / newton, 4),
            "pcg_iters": round(pcg / newton, 1),
            "solve_ms_per_newton_step": round(solve_ms / newton, 3),
            "newton_steps": newton, "block_jacobi_fallbacks": fallbacks,
            "same_system": {"block_jacobi_iters": same["bj_iter"],
                            "learned_iters": same["lpc_iter"],
                            "block_jacobi_solve_ms": round(same["bj_ms"], 3),
                            "learned_solve_ms": round(same["lpc_ms"], 3)},
            "training": {"systems": "the last Newton system of each of 12 time steps, "
                                    "stacked drape 8 x 64x64, seeds 0-2",
                         "adam_steps": args.lpc_steps,
                         "loss_first_last": [round(float(np.mean(hist[:20])), 4),
                                             round(float(np.mean(hist[-20:])), 4)],
                         "collect_s": round(t_collect, 2), "train_s": round(t_train, 2)}}


def walk_stats_step(one_step):
    """One more time step (after the timed window, untimed) with the BVH walk
    statistics on: per kind of walk, launches, queries, nodes taken off the
    walk stacks per query, CCD sweeps, the mean launch time (HIP events), and
    the walk's algorithmic bytes per launch -- the 64-byte walk record of
    every internal node entered (both children's boxes, tags and codes) --
    over that time."""
    from ppf_amd import _lib
    _lib.walk_stats_enable(True)
    try:
        one_step()
    finally:
        st = _lib.walk_stats()
        _lib.walk_stats_enable(False)
    out = {}
    for kind, w in st.items():
        if not w["launches"]:
            continue
        per = w["launches"]
        walk_bytes = 64 * (w["nodes"] - w["leaves"]) / per
        ms = w["ms"] / per
        out[kind] = {"launches": per, "queries_per_launch": round(w["queries"] / per),
                     "nodes_per_query": round(w["nodes"] / max(w["queries"], 1), 1),
                     "sweeps_per_launch": round(w["sweeps"] / per),
                     "ms_per_launch": round(ms, 3),
                     "walk_bytes_per_launch": round(walk_bytes),
                     "walk_gbs": round(walk_bytes / (ms * 1e6), 1) if ms > 0 else None}
    return out


def cpu_drape(args, start):
    """The oracle's first Newton iteration of the timed window (time step W+1),
    from the GPU's state at the start of the window: a bounded sample of the
    same workload (one Newton step, ~30 s on 16 host threads; a whole time
    step of 4 Newton steps takes ~110 s)."""
    import numpy as np

    from ppf_amd import scenes
    oracle, threads = oracle_module()
    sc = scenes.stacked_drape_scene(res=args.res, layers=args.layers)
    ns = args.layers * args.res * args.res
    sc["mesh"]["x"][:] = start["x"]
    sc["prev"][:] = start["prev"]
    sc["fix_delta"] = np.zeros_like(sc["fix_pos"])
    now = scenes.spin_fix_positions(sc, start["clock"])[ns:].copy()
    aim = scenes.spin_fix_positions(sc, start["clock"] + scenes.step_params()["dt"])[ns:].copy()
    sc["fix_pos"][ns:] = aim
    sc["fix_delta"][ns:] = aim - now
    st = oracle.OracleStepper(sc)
    p = scenes.step_params()
    p["prev_dt"] = start["prev_dt"]
    # exactly one Newton iteration: the loop stops at the second (status 5)
    # before the minimum count would let it finish the step
    p["max_newton_steps"] = 2
    p["min_newton_steps"] = 1000
    t0 = time.perf_counter()
    r = st.advance(p)
    dt = time.perf_counter() - t0
    if r["newton_steps"] != 1:
        raise RuntimeError(f"cpu sample: expected one Newton iteration, got {r}")
    return {"value": round(dt * 1e3, 1), "unit": "ms/Newton step",
            "cores": threads, "kind": "port", "host": host_cpu(),
            "sample": f"the first Newton iteration of time step {args.warmup + 1} of the same "
                      f"scene (the first step of the timed window, from the GPU's state): "
                      f"{r['pcg_iter_total']} PCG iterations, {r['pairs_max']} contact pairs, "
                      f"{dt:.1f} s; reference-algorithm CPU restatement (OpenMP, {threads} "
                      "threads), not the reference binary"}


def run_grid(args, dev):
    import numpy as np
    import torch

    from ppf_amd import scenes
    from ppf_amd.solver import BlockJacobiPCG

    res = args.res if args.res != 256 else 1024
    s = scenes.grid_spd_system(res)
    n, nnz = s["n"], len(s["col_index"])
    t = {k: torch.from_numpy(np.ascontiguousarray(s[k])).to(dev).reshape(-1)
         for k in ("row_offset", "col_index", "value", "cdiag", "b")}
    plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
    x = torch.zeros_like(t["b"])

    def step():
        x.zero_()
        return plan.solve(t["value"], t["cdiag"], t["b"], x, tol=1e-3)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = []
    for _ in range(args.steps):
        r = step()
        iters.append(r.iter)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    if r.status != 0:
        raise RuntimeError(f"PCG did not converge: status {r.status}")
    roof = spmv_roofline(plan, t["value"], t["cdiag"], t["b"], n, nnz,
                         "pcg spmv (k_spmv_ell<true,3> on this grid)",
                         traffic_from_profiles("pcg_spmv_grid1024.md", n, nnz))
    plan.close()
    workload = {"workload": f"kernel-isolation PCG: {res}x{res} triangulated grid, {n} rows "
                            f"of 3x3 fp32 blocks ({nnz} stored upper blocks), block-Jacobi "
                            "to 1e-3 L1 relative residual",
                "rows": n, "stored_blocks": nnz, "tol": 1e-3}
    out = head(args, ms, ms, workload, {"pcg_iters": int(np.median(iters)), "roofline": roof,
                                        "cpu_baseline": None})
    if not args.no_cpu:
        oracle, threads = oracle_module()
        t0 = time.perf_counter()
        ro = oracle.pcg(s, tol=1e-3)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(dt * 1e3, 1), "unit": "ms/solve",
                               "cores": threads, "kind": "port", "host": host_cpu(),
                               "sample": f"one full solve of the same system ({ro['iter']} "
                                         "iterations); reference-algorithm CPU restatement, "
                                         "not the reference binary"}
    return out


def run_batch(args, dev):
    """Many uneven scenes stepped together: 64 draped sheets (side
    log-uniform in 32..384 vertices, sphere radius uniform in 0.1..0.4 m,
    seed 5), all packed into one mesh with per-vertex scene ids and advanced
    by one many-scene stepper (every stage one launch over all scenes, each
    scene with its own Newton loop, solver scalars and line-search minimum),
    against the same 64 scenes advanced one stepper at a time. W untimed time
    steps, then K timed ones; a step is one time step of all 64 scenes."""
    import numpy as np
    import torch

    from ppf_amd import scenes
    from ppf_amd.stepper import Stepper, scene_to_device

    rng = np.random.default_rng(5)
    count = 64
    sides = np.exp(rng.uniform(np.log(32), np.log(384), count)).astype(int)
    radii = rng.uniform(0.1, 0.4, count)
    scs = [scenes.drape_scene(res=int(side), radius=float(rad), height=0.002)
           for side, rad in zip(sides, radii)]
    p = scenes.step_params()

    # one stepper per scene
    singles = []
    for sc in scs:
        t = scene_to_device(sc, dev)
        singles.append((t, Stepper(t), [p["dt"]]))

    def step_singles():
        out = []
        for t, st, pd in singles:
            q = dict(p)
            q["prev_dt"] = pd[0]
            r = st.advance(q)
            if r["status"] != 0:
                raise RuntimeError(f"single scene step failed: {r}")
            pd[0] = r["dt"]
            out.append(r)
        return out

    packed, rows = scenes.pack_scenes(scs)
    tb = scene_to_device(packed, dev)
    batch = Stepper(tb)
    prev_dt = [p["dt"]] * count

    def step_batch():
        res = batch.advance_scenes(p, prev_dt)
        for c, r in enumerate(res):
            if r["status"] != 0:
                raise RuntimeError(f"batched scene {c} step failed: {r}")
            prev_dt[c] = r["dt"]
        return res

    def timed(fn, label):
        for k in range(args.warmup):
            fn()
            progress(f"{label} warm-up step {k + 1}")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        newton = pcg = 0
        for k in range(args.steps):
            res = fn()
            newton += sum(r["newton_steps"] for r in res)
            pcg += sum(r["pcg_iter_total"] for r in res)
            progress(f"{label} step {k + 1}: newton {newton} pcg {pcg}")
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps, newton, pcg

    seq_ms, seq_newton, seq_pcg = timed(step_singles, "one stepper per scene")
    bat_ms, bat_newton, bat_pcg = timed(step_batch, "batched")
    xb = tb["mesh"]["x"].cpu().numpy().reshape(-1, 3)
    gap = 0.0
    for c, (t, st, _) in enumerate(singles):
        xs = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
        xc = xb[rows[c]:rows[c + 1]]
        gap = max(gap, float(np.abs(xs - xc).max() / np.abs(xs).max()))
        st.close()
    batch.close()
    n = int(rows[-1])
    workload = {"workload": f"{count} draped sheets, side log-uniform 32..384 vertices "
                            "(seed 5), sphere radius 0.1..0.4 m, 2 mm above it, gap 1e-3, "
                            "friction 0.5, dt 1/60; time steps "
                            f"{args.warmup + 1}..{args.warmup + args.steps} of all scenes",
                "scenes": count, "vertices": n,
                "vertices_min_max": [int(np.diff(rows).min()), int(np.diff(rows).max())]}
    return {"metric": "ms per time step of 64 uneven scenes, one many-scene stepper",
            "value": round(bat_ms, 4), "unit": "ms/time step", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(bat_ms, 4),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic", "config": workload,
            "one_stepper_per_scene_ms": round(seq_ms, 4),
            "speedup": round(seq_ms / bat_ms, 3),
            "newton_steps": {"batched": bat_newton, "one_per_scene": seq_newton},
            "pcg_iters": {"batched": bat_pcg, "one_per_scene": seq_pcg},
            "max_position_gap": gap, "roofline": None, "cpu_baseline": None}


def run_broad_narrow(args, dev):
    """Kernel isolation of the broad and narrow phase: 2,000,000 independent
    triangles (centroids uniform in a 1 m cube, vertices within 5e-3 m of
    them, seed 2), gap 1e-3. A step is one LBVH build of the face, edge and
    vertex trees plus the point-face / point-edge / point-point / edge-edge
    proximity enumeration (count pass keeping each query's first 8 pairs, scan,
    copy of the kept pairs, fill pass for queries with more)."""
    import numpy as np
    import torch

    from ppf_amd import scenes
    from ppf_amd.narrow import NarrowPhase, mesh_to_device

    nf = 2_000_000
    x, f = scenes.random_triangles(nf, seed=2)
    m = scenes.contact_mesh(x, f, ghat=1e-3)
    t = mesh_to_device(m, dev)
    nph = NarrowPhase()
    nph.build(t)
    pairs = nph.proximity(t)  # warm, sizes the pair buffer
    out = torch.empty_like(pairs) if pairs.shape[0] else torch.empty((1, 4), dtype=torch.int32,
                                                                     device=dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    build_ms = query_ms = 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev[0].record()
        nph.build(t)
        ev[1].record()
        got = nph.proximity(t, out)
        ev[2].record()
        torch.cuda.synchronize()
        build_ms += ev[0].elapsed_time(ev[1])
        query_ms += ev[1].elapsed_time(ev[2])
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    kinds = np.bincount(got[:, 0].cpu().numpy() & 255, minlength=4)
    nph.close()
    workload = {"workload": "broad + narrow phase kernel isolation: 2,000,000 independent "
                            "triangles, centroids U[0,1]^3, vertices within 5e-3 m (seed 2), "
                            "gap 1e-3; LBVH build of 3 trees + proximity enumeration",
                "faces": nf, "vertices": int(m["nv"]), "edges": int(m["ne"]),
                "pairs": int(got.shape[0]), "pairs_by_kind_pf_pe_pp_ee": kinds[:4].tolist()}
    out_line = {"metric": "ms per LBVH build + proximity enumeration (2M triangles)",
                "value": round(ms, 4), "unit": "ms/step", "n_gpus": 1, "steps": args.steps,
                "warmup": 1, "ms_per_step": round(ms, 4), "higher_is_better": False,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                "config": workload,
                "phase_ms": {"build": round(build_ms / args.steps, 3),
                             "proximity": round(query_ms / args.steps, 3)},
                "roofline": None, "cpu_baseline": None}
    if not args.no_cpu:
        oracle, threads = oracle_module()
        t0 = time.perf_counter()
        o = oracle.proximity(m)
        dt = time.perf_counter() - t0
        out_line["cpu_baseline"] = {"value": round(dt * 1e3, 1), "unit": "ms/step",
                                    "cores": threads, "kind": "port", "host": host_cpu(),
                                    "sample": f"the same 2M triangles, trees and enumeration "
                                              f"({len(o)} pairs); reference-algorithm CPU "
                                              "restatement, not the reference binary"}
    return out_line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--workload", choices=("drape-stack", "pcg-grid", "batch", "broad-narrow"),
                    default="drape-stack")
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--layers", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precond", choices=("bj", "both"), default="both",
                    help="both: also train and time the learned preconditioner")
    ap.add_argument("--lpc-steps", type=int, default=1600)
    ap.add_argument("--copy-calib", action="store_true",
                    help="first run a 256 MiB device copy (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device("cuda:0")
    if args.copy_calib:
        # A plain copy of known size: 256 MiB read + 256 MiB written. Its
        # FETCH_SIZE / WRITE_SIZE calibrate the counters of the kernels after.
        a = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    run = {"drape-stack": run_drape, "pcg-grid": run_grid, "batch": run_batch,
           "broad-narrow": run_broad_narrow}[args.workload]
    out = run(args, dev)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
