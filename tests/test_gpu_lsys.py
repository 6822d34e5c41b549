"""LSYS snapshots of the stepper's own Newton systems: a system written in
the reference dumper's format and read back solves like the original
(block Jacobi on the device): iteration counts within 3 (the diagonal blocks
are folded together on writing, so their inverses differ in the last bits)
and the read-back solution meets the solve tolerance on the original
operator."""
import numpy as np
import pytest
import torch

import oracle
from ppf_amd import lsys, scenes
from ppf_amd.solver import BlockJacobiPCG
from ppf_amd.stepper import Stepper, scene_to_device

pytestmark = pytest.mark.gpu


def test_stepper_system_round_trips_through_lsys(dev, tmp_path):
    sc = scenes.drape_scene(res=24, height=0.002)
    t = scene_to_device(sc, dev)
    st = Stepper(t)
    p = scenes.step_params()
    for _ in range(4):
        r = st.advance(p)
        assert r["status"] == 0
        p["prev_dt"] = r["dt"]
    s = st.system()
    st.close()
    path = tmp_path / "linsys_3.bin"
    lsys.write(path, s)
    back = lsys.read(path)
    out = []
    for sysd in (s, back):
        d = {k: torch.as_tensor(np.ascontiguousarray(np.asarray(
            sysd[k].cpu() if hasattr(sysd[k], "cpu") else sysd[k]).view(
            np.int32 if k in ("row_offset", "col_index") else np.float32).reshape(-1))).to(dev)
            for k in ("row_offset", "col_index", "value", "cdiag", "b")}
        plan = BlockJacobiPCG(d["row_offset"], d["col_index"])
        x = torch.zeros_like(d["b"])
        res = plan.solve(d["value"], d["cdiag"], d["b"], x, tol=1e-3)
        plan.close()
        out.append((res.iter, x.cpu().numpy()))
    host = {k: (s[k].cpu().numpy() if hasattr(s[k], "cpu") else s[k]) for k in s}
    host = {"n": len(host["row_offset"]) - 1, "row_offset": host["row_offset"].view(np.uint32),
            "col_index": host["col_index"].view(np.uint32), "value": host["value"],
            "cdiag": host["cdiag"], "b": host["b"]}
    rel = []
    for _, x in out:
        rres = host["b"] - oracle.spmv(host, x)
        rel.append(np.abs(rres).sum() / np.abs(host["b"]).sum())
    print(f"iterations {out[0][0]} / {out[1][0]} (original / read back), "
          f"relative L1 residual {rel[0]:.2e} / {rel[1]:.2e}")
    assert abs(out[0][0] - out[1][0]) <= 3
    assert rel[1] < 2e-3
