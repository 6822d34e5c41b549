"""Training and evaluation of the learned preconditioner on Newton systems the
solver itself produces (the (Hessian, residual) pairs of its Newton
iterations, as a linear-system dump records them). All compute is on the
device, through the library's own forward, backward and Adam kernels.
"""
import os
import time

import numpy as np
import torch

from . import lsys, scenes
from .learned import LearnedPreconditioner, init_weights
from .solver import BlockJacobiPCG
from .stepper import Stepper, scene_to_device


def collect_drape_systems(dev, res=64, layers=8, seeds=(0, 1, 2), steps=16, every=1,
                          lsys_dir=None):
    """Run the stacked drape (block-Jacobi solver) for `steps` time steps per
    seed and keep the last Newton system of every `every`-th step (device
    tensors: row_offset, col_index, value, cdiag, b). With `lsys_dir`, each
    kept system is also written there as an LSYS snapshot
    (linsys_<seed>_<step>.bin, the reference dumper's format)."""
    out = []
    for seed in seeds:
        sc = scenes.stacked_drape_scene(res=res, layers=layers, seed=seed)
        ns = layers * res * res
        t = scene_to_device(sc, dev)
        t["fix_delta"] = torch.zeros_like(t["mesh"]["x"])
        st = Stepper(t)
        p = scenes.step_params()
        clock = 0.0
        for k in range(steps):
            now = scenes.spin_fix_positions(sc, clock)[ns:].copy()
            aim = scenes.spin_fix_positions(sc, clock + p["dt"])[ns:].copy()
            t["fix_pos"].view(-1, 3)[ns:] = torch.from_numpy(aim).to(dev)
            t["fix_delta"].view(-1, 3)[ns:] = torch.from_numpy(aim - now).to(dev)
            r = st.advance(p)
            if r["status"] != 0:
                break
            p["prev_dt"] = r["dt"]
            clock += r["dt"]
            if k % every == 0:
                out.append(st.system())
                if lsys_dir is not None:
                    lsys.write(os.path.join(lsys_dir, f"linsys_{seed}_{k}.bin"), out[-1])
        st.close()
    return out


def train(systems, steps=400, lr=2e-3, nprobe=4, seed=0, weights=None, log=None):
    """Adam on the log Kaporin loss over the systems (round robin). Returns
    the preconditioner and the loss history."""
    dev = systems[0]["b"].device
    lpc = LearnedPreconditioner(init_weights(seed) if weights is None else weights)
    plans = [BlockJacobiPCG(s["row_offset"], s["col_index"]) for s in systems]
    gen = torch.Generator(device=dev).manual_seed(seed)
    hist = []
    for step in range(1, steps + 1):
        k = (step - 1) % len(systems)
        s, plan = systems[k], plans[k]
        n = s["b"].numel() // 3
        lpc.factor(plan, s["value"], s["cdiag"], s["b"])
        z = (torch.randint(0, 2, (nprobe, n, 3), generator=gen, device=dev).float() * 2 - 1)
        loss, _ = lpc.gradient(plan, s["value"], s["cdiag"], z.contiguous())
        lpc.adam(lr, step)
        hist.append(loss)
        if log and step % 100 == 0:
            log(f"lpc step {step}: loss {np.mean(hist[-100:]):.4f}")
    for p in plans:
        p.close()
    return lpc, hist


def compare(lpc, systems, tol=1e-3):
    """PCG iterations and solve time (ms, including the network for the
    learned factor) per system: block Jacobi vs the learned factor."""
    rows = []
    for s in systems:
        plan = BlockJacobiPCG(s["row_offset"], s["col_index"])
        x = torch.zeros_like(s["b"])
        plan.solve(s["value"], s["cdiag"], s["b"], x, tol=tol)  # graph capture warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x.zero_()
        bj = plan.solve(s["value"], s["cdiag"], s["b"], x, tol=tol)
        torch.cuda.synchronize()
        t_bj = (time.perf_counter() - t0) * 1e3
        lpc.factor(plan, s["value"], s["cdiag"], s["b"])
        x2 = torch.zeros_like(s["b"])
        lpc.solve(plan, s["value"], s["cdiag"], s["b"], x2, tol=tol)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lpc.factor(plan, s["value"], s["cdiag"], s["b"])
        x2.zero_()
        r = lpc.solve(plan, s["value"], s["cdiag"], s["b"], x2, tol=tol)
        torch.cuda.synchronize()
        t_l = (time.perf_counter() - t0) * 1e3
        rows.append(dict(bj_iter=bj.iter, bj_status=bj.status, bj_ms=t_bj, lpc_iter=r.iter,
                         lpc_status=r.status, lpc_ms=t_l))
        plan.close()
    return rows
