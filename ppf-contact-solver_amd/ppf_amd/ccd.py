"""CCD step filter over torch device buffers.

Mirrors contact::line_search (crates/ppf-cts-solver/src/cpp/contact/contact.cu:
2472-2740) and the ACCD kinds of contact/accd.hpp:547-696. All compute is in
libppf_mi355x.so.
"""
import ctypes

import torch

from . import _lib
from ._lib import check
from .narrow import _ptr, _stream, mesh_struct

vp = ctypes.c_void_p
u32 = ctypes.c_uint32
f32 = ctypes.c_float

STATIC_FLOAT = ("x", "v_ghat", "v_off", "e_ghat", "e_off", "f_ghat", "f_off")
STATIC_FRIC = ("v_fric", "e_fric", "f_fric")
STATIC_INDEX = ("edge", "face")
KIND_COLUMNS = {0: 4, 1: 3, 2: 2, 3: 4}


class StaticMeshC(ctypes.Structure):
    """ppf_static_mesh (include/ppf_mi355x.h)."""
    _fields_ = [("nv", u32), ("ne", u32), ("nf", u32), ("x", vp), ("edge", vp),
                ("face", vp), ("v_ghat", vp), ("v_off", vp), ("e_ghat", vp),
                ("e_off", vp), ("f_ghat", vp), ("f_off", vp), ("v_fric", vp),
                ("e_fric", vp), ("f_fric", vp), ("vscene", vp)]


class LsParamsC(ctypes.Structure):
    """ppf_ls_params (include/ppf_mi355x.h)."""
    _fields_ = [("x1", vp), ("vmass", vp), ("emass", vp), ("fmass", vp),
                ("max_t", f32), ("ccd_eps", f32), ("spheres", vp), ("nsphere", u32),
                ("floors", vp), ("nfloor", u32)]


def static_to_device(sm, dev):
    """Upload a static collision mesh dict (numpy) as device tensors."""
    import numpy as np
    out = {"nv": int(sm["nv"]), "ne": int(sm["ne"]), "nf": int(sm["nf"])}
    for k in STATIC_FLOAT:
        out[k] = torch.from_numpy(np.ascontiguousarray(sm[k], np.float32)).reshape(-1).to(dev)
    for k in STATIC_FRIC:
        if k in sm:
            out[k] = torch.from_numpy(np.ascontiguousarray(sm[k], np.float32)).reshape(-1).to(dev)
    for k in STATIC_INDEX:
        a = np.ascontiguousarray(sm[k], np.uint32).view(np.int32)
        out[k] = torch.from_numpy(a).reshape(-1).to(dev)
    if sm.get("vscene") is not None:
        a = np.ascontiguousarray(sm["vscene"], np.uint32).view(np.int32)
        out["vscene"] = torch.from_numpy(a).reshape(-1).to(dev)
    return out


def static_struct(t):
    if t is None:
        return StaticMeshC(nv=0, ne=0, nf=0)
    fields = {"nv": t["nv"], "ne": t["ne"], "nf": t["nf"]}
    for k in STATIC_FLOAT + STATIC_INDEX + STATIC_FRIC:
        if k not in t:
            continue
        ten = t[k]
        if not ten.is_cuda or not ten.is_contiguous():
            raise ValueError(f"{k}: expected a contiguous cuda tensor")
        fields[k] = ten.data_ptr() if ten.numel() else None
    if t.get("vscene") is not None:
        fields["vscene"] = t["vscene"].data_ptr()
    return StaticMeshC(**fields)


class LineSearch:
    """Largest certified fraction of a step x0 -> x1 (the contact CCD filter)."""

    def __init__(self):
        self.lib = _lib.load()
        h = vp()
        check(self.lib.ppf_ccd_create(ctypes.byref(h)), "ppf_ccd_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.lib.ppf_ccd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __call__(self, t, x1, vmass, emass, fmass, static=None, max_t=1.25,
                 ccd_eps=1e-7, rebuild=True):
        """t: device contact mesh at x0; x1, masses: device float tensors;
        static: device static mesh dict or None. Returns (toi, overlap)."""
        s = mesh_struct(t)
        ss = static_struct(static)
        for name, ten in (("x1", x1), ("vmass", vmass), ("emass", emass),
                          ("fmass", fmass)):
            if not ten.is_cuda or not ten.is_contiguous() or ten.dtype != torch.float32:
                raise ValueError(f"{name}: expected a contiguous float32 cuda tensor")
        p = LsParamsC(x1=_ptr(x1), vmass=_ptr(vmass), emass=_ptr(emass), fmass=_ptr(fmass),
                      max_t=max_t, ccd_eps=ccd_eps)
        toi, ov = f32(), ctypes.c_int()
        check(self.lib.ppf_line_search(self._h, ctypes.byref(s), ctypes.byref(ss),
                                       ctypes.byref(p), 1 if rebuild else 0,
                                       ctypes.byref(toi), ctypes.byref(ov), _stream()),
              "ppf_line_search")
        return float(toi.value), bool(ov.value)


def ccd_batch(kind, x0, x1, offset, ghat, max_t=1.25, ccd_eps=1e-7):
    """ACCD of n pairs of one kind: x0/x1 (n, columns, 3) float32 device
    tensors, offset/ghat (n,) float32. Returns (toi, overlap) device tensors."""
    cols = KIND_COLUMNS[kind]
    n = x0.shape[0]
    if x0.shape != (n, cols, 3) or x1.shape != (n, cols, 3):
        raise ValueError(f"x0/x1 must be ({n}, {cols}, 3)")
    if offset.shape != (n,) or ghat.shape != (n,):
        raise ValueError("offset/ghat must be (n,)")
    for ten in (x0, x1, offset, ghat):
        if not ten.is_cuda or not ten.is_contiguous() or ten.dtype != torch.float32:
            raise ValueError("expected contiguous float32 cuda tensors")
    toi = torch.empty(n, dtype=torch.float32, device=x0.device)
    ov = torch.empty(n, dtype=torch.int32, device=x0.device)
    lib = _lib.load()
    check(lib.ppf_ccd_batch(kind, n, _ptr(x0), _ptr(x1), _ptr(offset), _ptr(ghat),
                            max_t, ccd_eps, _ptr(toi), _ptr(ov), _stream()),
          "ppf_ccd_batch")
    return toi, ov
