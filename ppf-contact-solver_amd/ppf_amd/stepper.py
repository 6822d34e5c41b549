"""One time step of the contact solver over torch device buffers.

Mirrors main_helper::advance (crates/ppf-cts-solver/src/cpp/main/main.cu:
267-1649) for shells with fix pins, self contact and a static collision
mesh. All compute is in libppf_mi355x.so.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from .ccd import StaticMeshC, static_struct, static_to_device
from .elastic import SHELL_F32, SHELL_OPT_F32, SHELL_U32, ShellC
from .narrow import ContactMeshC, _ptr, _stream, mesh_struct, mesh_to_device

vp = ctypes.c_void_p
u32 = ctypes.c_uint32
f32 = ctypes.c_float
i32 = ctypes.c_int32


class SphereC(ctypes.Structure):
    """ppf_sphere (include/ppf_mi355x.h; data.rs:689-698)."""
    _fields_ = [("center", f32 * 3), ("ghat", f32), ("friction", f32), ("radius", f32),
                ("thickness", f32), ("bowl", ctypes.c_bool), ("reverse", ctypes.c_bool),
                ("kinematic", ctypes.c_bool)]


class FloorC(ctypes.Structure):
    """ppf_floor (include/ppf_mi355x.h; data.rs:702-709)."""
    _fields_ = [("ground", f32 * 3), ("ghat", f32), ("friction", f32), ("thickness", f32),
                ("up", f32 * 3), ("kinematic", ctypes.c_bool)]


def collider_arrays(t):
    """ctypes arrays of the analytic colliders in t['spheres'] / t['floors']
    (lists of dicts with the Sphere / Floor fields)."""
    sp = t.get("spheres") or []
    fl = t.get("floors") or []
    a = (SphereC * max(1, len(sp)))()
    for k, d in enumerate(sp):
        a[k] = SphereC(center=(f32 * 3)(*d["center"]), ghat=d["ghat"], friction=d["friction"],
                       radius=d["radius"], thickness=d.get("thickness", 0.0),
                       bowl=bool(d.get("bowl", False)), reverse=bool(d.get("reverse", False)),
                       kinematic=bool(d.get("kinematic", False)))
    b = (FloorC * max(1, len(fl)))()
    for k, d in enumerate(fl):
        b[k] = FloorC(ground=(f32 * 3)(*d["ground"]), ghat=d["ghat"], friction=d["friction"],
                      thickness=d.get("thickness", 0.0), up=(f32 * 3)(*d.get("up", (0, 1, 0))),
                      kinematic=bool(d.get("kinematic", False)))
    return a, len(sp), b, len(fl)


class SceneC(ctypes.Structure):
    """ppf_scene (include/ppf_mi355x.h)."""
    _fields_ = [("mesh", ContactMeshC), ("shell", ShellC), ("collision", StaticMeshC),
                ("prev", vp), ("emass", vp), ("v_fric", vp), ("e_fric", vp), ("f_fric", vp),
                ("fix_pos", vp), ("fix_delta", vp), ("spheres", vp), ("nsphere", u32),
                ("floors", vp), ("nfloor", u32)]


class StepParamsC(ctypes.Structure):
    """ppf_step_params (include/ppf_mi355x.h)."""
    _fields_ = [("dt", f32), ("prev_dt", f32), ("gravity", f32 * 3), ("cg_tol", f32),
                ("cg_max_iter", u32), ("target_toi", f32), ("line_search_max_t", f32),
                ("ccd_eps", f32), ("max_dx", f32), ("friction_eps", f32),
                ("friction_mode", i32), ("min_newton_steps", u32), ("max_newton_steps", u32),
                ("air_density", f32), ("air_friction", f32), ("wind", f32 * 3),
                ("disable_contact", i32), ("check_intersection", i32),
                ("constraint_tol", f32), ("eiganalysis_eps", f32)]


class StepResultC(ctypes.Structure):
    """ppf_step_result (include/ppf_mi355x.h)."""
    _fields_ = [("status", i32), ("newton_steps", u32), ("pcg_iter_total", u32),
                ("pcg_iter_max", u32), ("pairs_max", u32), ("intersections", u32),
                ("toi_advanced", f32), ("dt", f32), ("ms_elastic", f32), ("ms_contact", f32),
                ("ms_solve", f32), ("ms_line_search", f32), ("ms_total", f32),
                ("lpc_fallbacks", u32), ("device_alloc", u32), ("device_free", u32)]


def params_struct(p):
    kw = {k: v for k, v in p.items() if k in dict(StepParamsC._fields_)}
    kw["gravity"] = (f32 * 3)(*p["gravity"])
    kw["wind"] = (f32 * 3)(*p["wind"])
    return StepParamsC(**kw)


def result_dict(r):
    return {k: getattr(r, k) for k, _ in StepResultC._fields_}


def scene_to_device(sc, dev):
    """Upload a scene dict (scenes.drape_scene layout) as device tensors."""
    t = {"mesh": mesh_to_device(sc["mesh"], dev)}
    sh = {k: int(sc["shell"][k]) for k in ("nv", "nf", "nh")}
    for k in SHELL_F32:
        if k in ("x", "x0", "target"):
            continue
        sh[k] = torch.from_numpy(np.ascontiguousarray(sc["shell"][k], np.float32).reshape(-1)).to(dev)
    for k in SHELL_U32:
        a = np.ascontiguousarray(sc["shell"][k], np.uint32).reshape(-1).view(np.int32)
        sh[k] = torch.from_numpy(a).to(dev)
    for k in SHELL_OPT_F32:
        if sc["shell"].get(k) is not None:
            sh[k] = torch.from_numpy(np.ascontiguousarray(sc["shell"][k], np.float32)
                                     .reshape(-1)).to(dev)
    t["shell"] = sh
    t["static"] = static_to_device(sc["static"], dev) if sc.get("static") else None
    f = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32).reshape(-1)).to(dev)
    t["prev"] = f(sc["prev"])
    t["emass"] = f(sc["emass"])
    t["fric"] = {k: f(v) for k, v in sc["fric"].items()}
    t["fix_pos"] = f(sc["fix_pos"]) if sc.get("fix_pos") is not None else None
    t["fix_delta"] = f(sc["fix_delta"]) if sc.get("fix_delta") is not None else None
    t["spheres"] = list(sc.get("spheres") or [])
    t["floors"] = list(sc.get("floors") or [])
    return t


def scene_struct(t):
    sh = t["shell"]
    fields = {k: sh[k] for k in ("nv", "nf", "nh")}
    for k in SHELL_F32 + SHELL_U32 + SHELL_OPT_F32:
        if k in sh:
            fields[k] = sh[k].data_ptr() if sh[k].numel() else None
    st = static_struct(t["static"]) if t["static"] is not None else StaticMeshC(nv=0, ne=0, nf=0)
    return SceneC(mesh=mesh_struct(t["mesh"]), shell=ShellC(**fields), collision=st,
                  prev=_ptr(t["prev"]), emass=_ptr(t["emass"]), v_fric=_ptr(t["fric"]["v"]),
                  e_fric=_ptr(t["fric"]["e"]), f_fric=_ptr(t["fric"]["f"]),
                  fix_pos=_ptr(t["fix_pos"]),
                  fix_delta=_ptr(t.get("fix_delta")))


class Stepper:
    """Time stepping of a device scene (mesh['x'] and prev are advanced in
    place)."""

    def __init__(self, t):
        self.lib = _lib.load()
        self.t = t
        self._sc = scene_struct(t)
        h = vp()
        check(self.lib.ppf_stepper_create(ctypes.byref(self._sc), _stream(), ctypes.byref(h)),
              "ppf_stepper_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.lib.ppf_stepper_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_preconditioner(self, lpc):
        """Use a LearnedPreconditioner for the Newton solves (None: block Jacobi)."""
        self._lpc = lpc
        check(self.lib.ppf_stepper_set_preconditioner(self._h, lpc._h if lpc else None),
              "ppf_stepper_set_preconditioner")

    def solve_order(self):
        """Row order of the stepper's solves (int32 device tensor, new row k =
        vertex order[k])."""
        dev = self.t["mesh"]["x"].device
        n = self.t["mesh"]["x"].numel() // 3
        out = torch.empty(n, dtype=torch.int32, device=dev)
        check(self.lib.ppf_stepper_solve_order(self._h, _ptr(out), _stream()),
              "ppf_stepper_solve_order")
        return out

    def system(self):
        """Copy of the last Newton system solved: dict of device tensors
        row_offset (n+1, int32), col_index (nnz, int32), value (9 nnz),
        cdiag (9 n), b (3 n)."""
        n, nnz = u32(), u32()
        check(self.lib.ppf_stepper_system(self._h, ctypes.byref(n), ctypes.byref(nnz), None,
                                          None, None, None, None, _stream()),
              "ppf_stepper_system")
        dev = self.t["mesh"]["x"].device
        n, nnz = n.value, nnz.value
        out = {"row_offset": torch.empty(n + 1, dtype=torch.int32, device=dev),
               "col_index": torch.empty(nnz, dtype=torch.int32, device=dev),
               "value": torch.empty(9 * nnz, dtype=torch.float32, device=dev),
               "cdiag": torch.empty(9 * n, dtype=torch.float32, device=dev),
               "b": torch.empty(3 * n, dtype=torch.float32, device=dev)}
        check(self.lib.ppf_stepper_system(self._h, ctypes.byref(u32()), ctypes.byref(u32()),
                                          *(_ptr(out[k]) for k in ("row_offset", "col_index",
                                                                   "value", "cdiag", "b")),
                                          _stream()), "ppf_stepper_system")
        return out

    def advance_scenes(self, params, prev_dt):
        """One time step of every scene of a many-scene stepper; prev_dt is
        a sequence of one value per scene. Returns one result dict per scene."""
        ns = int(self.t["mesh"]["nscene"])
        p = params_struct(params)
        pd = (f32 * ns)(*[float(v) for v in prev_dt])
        r = (StepResultC * ns)()
        check(self.lib.ppf_stepper_advance_scenes(self._h, ctypes.byref(self._sc), ctypes.byref(p),
                                                  pd, r, _stream()), "ppf_stepper_advance_scenes")
        return [result_dict(x) for x in r]

    def _colliders(self):
        # the analytic colliders as they are now (kinematic ones move)
        self._coll = collider_arrays(self.t)
        sp, nsp, fl, nfl = self._coll
        self._sc.spheres = ctypes.cast(sp, vp)
        self._sc.nsphere = nsp
        self._sc.floors = ctypes.cast(fl, vp)
        self._sc.nfloor = nfl

    def advance(self, params):
        p = params_struct(params)
        self._colliders()
        r = StepResultC()
        check(self.lib.ppf_stepper_advance(self._h, ctypes.byref(self._sc), ctypes.byref(p),
                                           ctypes.byref(r), _stream()), "ppf_stepper_advance")
        return result_dict(r)
