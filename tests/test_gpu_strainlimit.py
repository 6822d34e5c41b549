"""GPU parity of strain limiting (through the C ABI) against the CPU oracle.

* The stage entries: the strain-limit terms use only +, -, *, /, sqrt, fabs
  and fmax (correctly rounded on both sides) in the oracle's order, and the
  line search is the same bisection: force, blocks and toi are bit-exact
  (bending off, so no atan2f precedes them).
* Stepping: a draped sheet with a strain limit and a soft hanging sheet whose
  pins pull it apart (the limit's line search then shortens steps) against
  the oracle's step from the same state, at the step tests' position bar;
  every face stays below its limit on the GPU.
"""
import copy

import numpy as np
import pytest
import torch

import oracle
from ppf_amd import scenes
from ppf_amd.elastic import ElasticAssembler, shell_to_device
from ppf_amd.stepper import Stepper, scene_to_device
from test_gpu_step import (SOLVE_BALL_FACTOR, _compare, solve_error_vs_fp64, ulp_nearest,
                           within_ulp_copies)
from test_oracle_strainlimit import max_stretch64, one_face

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("res,sl,pins", [(16, 0.1, 0), (64, 0.05, 5), (64, 0.02, 0)])
def test_strain_limit_terms_bit_exact(dev, res, sl, pins):
    sh = scenes.shell_scene(res, seed=6, pins=pins, stretch=0.08)
    sh["h_bend"][:] = 0.0
    sh["f_strainlimit"] = np.full(sh["nf"], sl, np.float32)
    sh["f_strainlimit"][::11] = 0.0       # some faces unlimited
    sh["fflag"][::23] = 1                 # some excluded
    t = shell_to_device(sh, dev)
    asm = ElasticAssembler(t)
    rp, ci = asm.pattern()
    force, cdiag, value = asm.assemble(t, sh["dt"])
    el_force, el_value = force.cpu().numpy(), value.cpu().numpy()
    asm.strain_limit(t, force, value, 1e-2)
    rp = rp.cpu().numpy().astype(np.uint32)
    ci = ci.cpu().numpy().astype(np.uint32)
    f, v = force.cpu().numpy(), value.cpu().numpy()[:9 * len(ci)]
    of, _, ov, miss = oracle.elastic_assemble(sh, rp, ci, sh["dt"])
    assert miss == 0
    assert np.array_equal(el_force.view(np.uint32), of.view(np.uint32))
    sf, sv = oracle.strainlimit_assemble(sh, sh["x"], sh["x0"], rp, ci, of, ov)
    changed = np.mean(v != el_value[:9 * len(ci)])
    print(f"res {res} sl {sl}: {changed:.3f} of block entries changed by the limit")
    assert changed > 0.05
    assert np.array_equal(f.view(np.uint32), sf.view(np.uint32))
    assert np.array_equal(v.view(np.uint32), sv.reshape(-1).view(np.uint32))
    asm.close()


@pytest.mark.parametrize("scale,sl,shrink", [(1.2, 0.05, None), (1.3, 0.05, 0.95),
                                             (1.02, 0.05, None)])
def test_strain_limit_toi_one_face_bit_exact(dev, scale, sl, shrink):
    sh, x0, x1 = one_face(scale, sl, shrink)
    full = dict(sh, nh=0, x=x0, x0=x0, target=x0, hinge=np.zeros(0, np.uint32),
                v2f_off=np.array([0, 1, 2, 3], np.uint32), v2f=np.zeros(3, np.uint32),
                vmass=np.ones(3, np.float32), varea=np.ones(3, np.float32),
                vflag=np.zeros(3, np.uint32), f_mu=np.ones(1, np.float32),
                f_lambda=np.ones(1, np.float32), h_rest=np.zeros(0, np.float32),
                h_length=np.zeros(0, np.float32), h_area=np.zeros(0, np.float32),
                h_bend=np.zeros(0, np.float32), hflag=np.zeros(0, np.uint32))
    if shrink is None:
        full.pop("f_shrink")
    t = shell_to_device(full, dev)
    asm = ElasticAssembler(t)
    g = asm.strain_limit_toi(t, torch.from_numpy(x0.reshape(-1)).to(dev),
                             torch.from_numpy(x1.reshape(-1)).to(dev), 1.25)
    o = oracle.strainlimit_toi(sh, x0, x1, 1.25)
    asm.close()
    assert np.float32(g) == np.float32(np.float32(o) / np.float32(1.25)), (g, o)


def test_strain_limit_toi_sheet_bit_exact(dev):
    sh = scenes.shell_scene(64, seed=8, stretch=0.02)
    sh["f_strainlimit"] = np.full(sh["nf"], 0.03, np.float32)
    rng = np.random.default_rng(2)
    x0 = sh["x0"].astype(np.float32)      # the smooth state: below the limit
    x1 = (x0 * np.float32(1.06) + rng.normal(0, 2e-4, x0.shape)).astype(np.float32)
    t = shell_to_device(sh, dev)
    asm = ElasticAssembler(t)
    g = asm.strain_limit_toi(t, torch.from_numpy(x0.reshape(-1)).to(dev),
                             torch.from_numpy(x1.reshape(-1)).to(dev), 1.25)
    o = oracle.strainlimit_toi(sh, x0, x1, 1.25)
    asm.close()
    print(f"toi {g:.7g} (oracle {o / 1.25:.7g})")
    assert g < 1.0
    assert np.float32(g) == np.float32(np.float32(o) / np.float32(1.25))


def test_strain_limited_drape_vs_oracle(dev):
    sc = scenes.drape_scene(res=16, height=0.003, strain_limit=0.01)
    _compare(dev, sc, 10, scenes.step_params())


def test_pulled_sheet_vs_oracle(dev):
    # A soft sheet (Young 300) hanging from two corners that move apart:
    # the strain limit's line search shortens several steps.
    res, sl, v = 8, 0.01, np.float32(0.004)
    pins = [0, res - 1]
    sc = scenes.drape_scene(res=res, height=0.5, young=300.0, pins=pins, strain_limit=sl)
    sc["fix_delta"] = np.zeros_like(sc["fix_pos"])
    sco = copy.deepcopy(sc)
    t = scene_to_device(sc, dev)
    gpu = Stepper(t)
    orc = oracle.OracleStepper(sco)
    p = scenes.step_params(disable_contact=1)
    faces, inv = sc["shell"]["face"], sc["shell"]["inv_rest"]
    shortened = 0
    for k in range(12):
        xg = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
        pg = t["prev"].cpu().numpy().reshape(-1, 3)
        aim = xg[pins] + np.array([[-v, 0, 0], [v, 0, 0]], np.float32)
        delta = np.array([[-v, 0, 0], [v, 0, 0]], np.float32)
        t["fix_pos"].view(-1, 3)[pins] = torch.from_numpy(aim).to(dev)
        t["fix_delta"].view(-1, 3)[pins] = torch.from_numpy(delta).to(dev)
        sco["mesh"]["x"][:] = xg
        sco["prev"][:] = pg
        sco["fix_pos"][pins] = aim
        sco["fix_delta"][pins] = delta
        orc.delta[:] = sco["fix_delta"].reshape(-1)
        fix0 = sco["fix_pos"].copy()
        rg = gpu.advance(p)
        ro = orc.advance(p)
        xn = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
        xo = sco["mesh"]["x"]
        scale = np.abs(xo).max()
        err = np.abs(xn - xo).max() / scale
        eg, eo = solve_error_vs_fp64(gpu.system(), p["cg_tol"], scale)
        bar = max(1e-5, SOLVE_BALL_FACTOR * max(eg, eo))
        near = ulp_nearest(orc, sco, xg, pg, p, xn, scale, fix0) if err >= bar else None
        stretch = max_stretch64(xn.astype(np.float64), faces, inv).max()
        line = (f"step {k}: status {rg['status']}/{ro['status']} newton "
                f"{rg['newton_steps']}/{ro['newton_steps']} toi {rg['toi_advanced']:.6f}/"
                f"{ro['toi_advanced']:.6f} pos err {err:.2e} bar {bar:.2e} max stretch "
                f"{stretch:.5f}" + ("" if near is None else
                                    f"; one-ulp copies: nearest {near[0]:.2e}, spread "
                                    f"{near[1]:.2e}"))
        print(line, flush=True)
        assert rg["status"] == ro["status"] == 0, line
        assert rg["newton_steps"] == ro["newton_steps"], line
        assert abs(rg["toi_advanced"] - ro["toi_advanced"]) <= 1e-3, line
        assert err < bar or within_ulp_copies(err, near, bar), line
        assert stretch < sl + 1e-5, line
        shortened += rg["toi_advanced"] < 1.0
        p["prev_dt"] = rg["dt"]
    gpu.close()
    assert shortened >= 2
