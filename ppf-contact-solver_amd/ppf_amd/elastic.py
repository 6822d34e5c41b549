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


class ShellC(ctypes.Structure):
    """ppf_shell (include/ppf_mi355x.h)."""
    _fields_ = [("nv", u32), ("nf", u32), ("nh", u32), ("x", vp), ("x0", vp),
                ("target", vp), ("face", vp), ("hinge", vp), ("v2f_off", vp), ("v2f", vp),
                ("vmass", vp), ("varea", vp), ("vflag", vp), ("fmass", vp),
                ("inv_rest", vp), ("f_mu", vp), ("f_lambda", vp), ("fflag", vp),
                ("h_rest", vp), ("h_length", vp), ("h_area", vp), ("h_bend", vp),
                ("hflag", vp)]


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
    return out


def shell_struct(t):
    fields = {k: t[k] for k in ("nv", "nf", "nh")}
    for k in SHELL_F32 + SHELL_U32:
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
