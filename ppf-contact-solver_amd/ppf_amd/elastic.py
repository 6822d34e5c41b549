"""Inertia, air drag, membrane and bending assembly over torch device buffers.

Mirrors energy::embed_momentum_force_hessian and the shell part of
energy::embed_elastic_force_hessian (crates/ppf-cts-solver/src/cpp/energy/
energy.cu:182-553, 700-794). All compute is in libppf_mi355x.so.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from .narrow import _copy_dev, _ptr, _stream

vp = ctypes.c_void_p
u32 = ctypes.c_uint32
f32 = ctypes.c_float

SHELL_U32 = ("face", "hinge", "v2f_off", "v2f", "vflag", "fflag", "hflag")
SHELL_F32 = ("x", "x0", "target", "vmass", "varea", "fmass", "inv_rest", "f_mu",
             "f_lambda", "h_rest", "h_length", "h_area", "h_bend")
# optional per-face / per-hinge arrays (absent: no strain limit, shrink 1, no
# plasticity)
SHELL_OPT_F32 = ("f_strainlimit", "f_shrink", "f_plasticity", "f_plastic_threshold",
                 "h_plasticity", "h_plastic_threshold")


class ShellC(ctypes.Structure):
    """ppf_shell (include/ppf_mi355x.h)."""
    _fields_ = [("nv", u32), ("nf", u32), ("nh", u32), ("x", vp), ("x0", vp),
                ("target", vp), ("face", vp), ("hinge", vp), ("v2f_off", vp), ("v2f", vp),
                ("vmass", vp), ("varea", vp), ("vflag", vp), ("fmass", vp),
                ("inv_rest", vp), ("f_mu", vp), ("f_lambda", vp), ("fflag", vp),
                ("h_rest", vp), ("h_length", vp), ("h_area", vp), ("h_bend", vp),
                ("hflag", vp), ("f_strainlimit", vp), ("f_shrink", vp),
                ("f_plasticity", vp), ("f_plastic_threshold", vp), ("h_plasticity", vp),
                ("h_plastic_threshold", vp)]


class ElasticParamsC(ctypes.Structure):
    """ppf_elastic_params (include/ppf_mi355x.h)."""
    _fields_ = [("dt", f32), ("air_density", f32), ("air_friction", f32),
                ("wind", f32 * 3), ("scene_dt", ctypes.c_void_p), ("vscene", ctypes.c_void_p)]


def shell_to_device(sh, dev):
    out = {k: int(sh[k]) for k in ("nv", "nf", "nh")}
    for k in SHELL_F32:
        out[k] = torch.from_numpy(np.ascontiguousarray(sh[k], np.float32).reshape(-1)).to(dev)
    for k in SHELL_U32:
        a = np.ascontiguousarray(sh[k], np.uint32).reshape(-1).view(np.int32)
        out[k] = torch.from_numpy(a).to(dev)
    for k in SHELL_OPT_F32:
        if sh.get(k) is not None:
            out[k] = torch.from_numpy(np.ascontiguousarray(sh[k], np.float32).reshape(-1)).to(dev)
    return out


def shell_struct(t):
    fields = {k: t[k] for k in ("nv", "nf", "nh")}
    for k in SHELL_F32 + SHELL_U32 + tuple(k for k in SHELL_OPT_F32 if t.get(k) is not None):
        ten = t[k]
        if not ten.is_cuda or not ten.is_contiguous():
            raise ValueError(f"{k}: expected a contiguous cuda tensor")
        fields[k] = ten.data_ptr() if ten.numel() else None
    return ShellC(**fields)


class ElasticAssembler:
    """Fixed upper pattern of a shell and its elastic + inertia assembly."""

    def __init__(self, t):
        self.lib = _lib.load()
        s = shell_struct(t)
        h = vp()
        check(self.lib.ppf_elastic_create(ctypes.byref(s), _stream(), ctypes.byref(h)),
              "ppf_elastic_create")
        self._h = h
        self.nv = t["nv"]
        self.dev = t["x"].device

    def close(self):
        if getattr(self, "_h", None):
            self.lib.ppf_elastic_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pattern(self):
        """(row_offset, col_index) as int32 device tensors (copies)."""
        rp, ci, nnz = vp(), vp(), u32()
        check(self.lib.ppf_elastic_pattern(self._h, ctypes.byref(rp), ctypes.byref(ci),
                                           ctypes.byref(nnz)), "ppf_elastic_pattern")
        return (_copy_dev(rp.value, self.nv + 1, torch.int32, self.dev),
                _copy_dev(ci.value, nnz.value, torch.int32, self.dev))

    def nnz(self):
        n = u32()
        check(self.lib.ppf_elastic_pattern(self._h, None, None, ctypes.byref(n)),
              "ppf_elastic_pattern")
        return n.value

    def assemble(self, t, dt, air_density=1e-3, air_friction=0.2, wind=(0.0, 0.0, 0.0),
                 out=None):
        """Returns (force (3nv), cdiag (9nv), value (9 nnz)) device tensors."""
        s = shell_struct(t)
        p = ElasticParamsC(dt=dt, air_density=air_density, air_friction=air_friction,
                           wind=(f32 * 3)(*wind))
        if out is None:
            out = (torch.empty(3 * self.nv, dtype=torch.float32, device=self.dev),
                   torch.empty(9 * self.nv, dtype=torch.float32, device=self.dev),
                   torch.empty(9 * max(1, self.nnz()), dtype=torch.float32, device=self.dev))
        force, cdiag, value = out
        check(self.lib.ppf_elastic_assemble(self._h, ctypes.byref(s), ctypes.byref(p),
                                            _ptr(force), _ptr(cdiag), _ptr(value), _stream()),
              "ppf_elastic_assemble")
        return force, cdiag, value

    def strain_limit(self, t, force, value, eiganalysis_eps=1e-2):
        """Adds the strain-limit terms at t['x'] (stiffness from t['x0'] and
        the elastic blocks in `value`) to force and value in place
        (strainlimiting.cu:46-100)."""
        s = shell_struct(t)
        check(self.lib.ppf_elastic_strain_limit(self._h, ctypes.byref(s), eiganalysis_eps,
                                                _ptr(force), _ptr(value), _stream()),
              "ppf_elastic_strain_limit")
        return force, value

    def plasticity(self, t, dt, inv_rest=None, h_rest=None):
        """Rest-shape creep at t['x'] (plasticity.cu:21-78, 145-185):
        inv_rest / h_rest (default t's own) are updated in place."""
        s = shell_struct(t)
        inv_rest = t["inv_rest"] if inv_rest is None else inv_rest
        h_rest = t["h_rest"] if h_rest is None else h_rest
        check(self.lib.ppf_elastic_plasticity(self._h, ctypes.byref(s), dt, _ptr(inv_rest),
                                              _ptr(h_rest) if h_rest.numel() else None,
                                              _stream()),
              "ppf_elastic_plasticity")
        return inv_rest, h_rest

    def strain_limit_toi(self, t, x0, x1, line_search_max_t=1.25):
        """Fraction of the step x0 -> x1 (in units of line_search_max_t) that
        keeps every limited face below its limit (strainlimiting.cu:243-305)."""
        s = shell_struct(t)
        out = f32()
        check(self.lib.ppf_elastic_strain_limit_toi(self._h, ctypes.byref(s), _ptr(x0), _ptr(x1),
                                                    line_search_max_t, ctypes.byref(out),
                                                    _stream()),
              "ppf_elastic_strain_limit_toi")
        return out.value
