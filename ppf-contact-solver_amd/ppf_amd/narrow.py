"""LBVH + contact proximity narrow phase over torch device buffers.

Mirrors the detection half of contact::embed_contact_force_hessian
(crates/ppf-cts-solver/src/cpp/contact/contact.cu:1398-1787) and the LBVH
builds of lbvh/lbvh.cu:863-1026. All compute is in libppf_mi355x.so.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check

vp = ctypes.c_void_p
u32 = ctypes.c_uint32

FLOAT_FIELDS = ("x", "v_ghat", "v_off", "e_ghat", "e_off", "f_ghat", "f_off")
INDEX_FIELDS = ("edge", "face", "v2e_off", "v2e", "v2f_off", "v2f", "e2f_off",
                "e2f", "vflag", "vbody", "eflag", "fflag")
PAIR_DTYPE = np.dtype([("kind", np.uint32), ("a", np.uint32), ("b", np.uint32),
                       ("count", np.uint32)])
# ppf_isect_record (include/ppf_mi355x.h)
ISECT_DTYPE = np.dtype([("type", np.uint32), ("elem0", np.uint32), ("elem1", np.uint32),
                        ("num_verts0", np.uint32), ("num_verts1", np.uint32),
                        ("positions", np.float32, (15,))])


class ContactMeshC(ctypes.Structure):
    """ppf_contact_mesh (include/ppf_mi355x.h)."""
    _fields_ = [("nv", u32), ("ne", u32), ("nf", u32), ("x", vp),
                ("edge", vp), ("face", vp),
                ("v2e_off", vp), ("v2e", vp), ("v2f_off", vp), ("v2f", vp),
                ("e2f_off", vp), ("e2f", vp),
                ("v_ghat", vp), ("v_off", vp), ("e_ghat", vp), ("e_off", vp),
                ("f_ghat", vp), ("f_off", vp),
                ("vflag", vp), ("vbody", vp), ("eflag", vp), ("fflag", vp),
                ("vscene", vp), ("nscene", u32)]


def mesh_to_device(m, dev):
    """Upload a contact-mesh dict (numpy) as contiguous device tensors."""
    out = {"nv": int(m["nv"]), "ne": int(m["ne"]), "nf": int(m["nf"])}
    for k in FLOAT_FIELDS:
        out[k] = torch.from_numpy(np.ascontiguousarray(m[k], np.float32)).reshape(-1).to(dev)
    for k in INDEX_FIELDS:
        a = np.ascontiguousarray(m[k], np.uint32).view(np.int32)
        out[k] = torch.from_numpy(a).reshape(-1).to(dev)
    if m.get("vscene") is not None:  # many scenes in one mesh
        a = np.ascontiguousarray(m["vscene"], np.uint32).view(np.int32)
        out["vscene"] = torch.from_numpy(a).reshape(-1).to(dev)
        out["nscene"] = int(m["nscene"])
    return out


def mesh_struct(t):
    fields = {"nv": t["nv"], "ne": t["ne"], "nf": t["nf"]}
    for k in FLOAT_FIELDS + INDEX_FIELDS:
        ten = t[k]
        if not ten.is_cuda or not ten.is_contiguous():
            raise ValueError(f"{k}: expected a contiguous cuda tensor")
        fields[k] = ten.data_ptr() if ten.numel() else None
    if t.get("vscene") is not None:
        fields["vscene"] = t["vscene"].data_ptr()
        fields["nscene"] = t["nscene"]
    return ContactMeshC(**fields)


def _stream():
    return vp(torch.cuda.current_stream().cuda_stream)


class NarrowPhase:
    def __init__(self):
        self.lib = _lib.load()
        h = vp()
        check(self.lib.ppf_narrow_create(ctypes.byref(h)), "ppf_narrow_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.lib.ppf_narrow_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self, t):
        s = mesh_struct(t)
        check(self.lib.ppf_narrow_build(self._h, ctypes.byref(s), _stream()),
              "ppf_narrow_build")

    def tree(self, kind):
        n, root = u32(), u32()
        check(self.lib.ppf_narrow_tree(self._h, kind, None, None, ctypes.byref(n),
                                       ctypes.byref(root), _stream()), "ppf_narrow_tree")
        nn = max(1, 2 * n.value - 1)
        dev = torch.device("cuda")
        nodes = torch.zeros(2 * nn, dtype=torch.int32, device=dev)
        boxes = torch.zeros(8 * nn, dtype=torch.float32, device=dev)
        check(self.lib.ppf_narrow_tree(self._h, kind, vp(nodes.data_ptr()),
                                       vp(boxes.data_ptr()), ctypes.byref(n),
                                       ctypes.byref(root), _stream()), "ppf_narrow_tree")
        return (nodes.cpu().numpy().view(np.uint32).reshape(-1, 2),
                boxes.cpu().numpy().reshape(-1, 8), root.value)

    def count(self, t):
        s = mesh_struct(t)
        c = ctypes.c_uint64()
        check(self.lib.ppf_narrow_proximity(self._h, ctypes.byref(s), None, 0,
                                            ctypes.byref(c), _stream()),
              "ppf_narrow_proximity")
        return c.value

    def proximity(self, t, out=None):
        """Pairs as a device tensor (k, 4) int32 (kind, a, b, count)."""
        s = mesh_struct(t)
        c = ctypes.c_uint64()
        if out is None:
            check(self.lib.ppf_narrow_proximity(self._h, ctypes.byref(s), None, 0,
                                                ctypes.byref(c), _stream()),
                  "ppf_narrow_proximity")
            out = torch.empty((max(1, c.value), 4), dtype=torch.int32,
                              device=t["x"].device)
        cap = out.shape[0]
        check(self.lib.ppf_narrow_proximity(self._h, ctypes.byref(s),
                                            vp(out.data_ptr()), cap, ctypes.byref(c),
                                            _stream()), "ppf_narrow_proximity")
        if c.value > cap:
            raise _lib.PPFError(f"pair buffer too small: {c.value} > {cap}")
        return out[:c.value]

    def static_pairs(self, t, static, vmass, emass, fmass):
        """Collision-mesh pairs (kinds 4-6) as a (k, 4) int32 device tensor."""
        from .ccd import static_struct
        s = mesh_struct(t)
        ss = static_struct(static)
        c = ctypes.c_uint64()
        args = (vp(vmass.data_ptr()), vp(emass.data_ptr()), vp(fmass.data_ptr()))
        check(self.lib.ppf_narrow_static(self._h, ctypes.byref(s), ctypes.byref(ss), *args,
                                         None, 0, ctypes.byref(c), _stream()),
              "ppf_narrow_static")
        out = torch.empty((max(1, c.value), 4), dtype=torch.int32, device=t["x"].device)
        check(self.lib.ppf_narrow_static(self._h, ctypes.byref(s), ctypes.byref(ss), *args,
                                         vp(out.data_ptr()), out.shape[0], ctypes.byref(c),
                                         _stream()), "ppf_narrow_static")
        return out[:c.value]

    def intersection_check(self, t, vmass, emass, fmass, static=None, surface=None,
                           cap=256):
        """Intersection check at t['x'] (build() at the same positions first):
        (count of intersecting pairs, records[:min(count, cap)])."""
        from .ccd import static_struct
        s = mesh_struct(t)
        ss = static_struct(static)
        for name, ten in (("vmass", vmass), ("emass", emass), ("fmass", fmass)):
            if not ten.is_cuda or not ten.is_contiguous() or ten.dtype != torch.float32:
                raise ValueError(f"{name}: expected a contiguous float32 cuda tensor")
        rec = np.zeros(max(1, cap), ISECT_DTYPE)
        n = u32()
        nsurf = t["nv"] if surface is None else int(surface)
        check(self.lib.ppf_intersection_check(
            self._h, ctypes.byref(s), ctypes.byref(ss), vp(vmass.data_ptr()),
            vp(emass.data_ptr()), vp(fmass.data_ptr()), nsurf,
            rec.ctypes.data_as(vp), cap, ctypes.byref(n), _stream()),
            "ppf_intersection_check")
        return n.value, rec[:min(n.value, cap)]


def pairs_to_numpy(t):
    a = t.cpu().numpy().astype(np.uint32)
    out = np.zeros(len(a), PAIR_DTYPE)
    for i, k in enumerate(("kind", "a", "b", "count")):
        out[k] = a[:, i]
    return out


class ContactParamsC(ctypes.Structure):
    """ppf_contact_params (include/ppf_mi355x.h)."""
    _fields_ = [("x0", vp), ("mass", vp), ("v_fric", vp), ("e_fric", vp),
                ("f_fric", vp), ("friction_mode", ctypes.c_int32),
                ("friction_eps", ctypes.c_float), ("el_rp", vp), ("el_col", vp),
                ("el_val", vp), ("sm", vp)]


def _ptr(t):
    return None if t is None else t.data_ptr()


class ContactAssembler:
    """Contact gradient + upper-triangle 3x3 blocks of a pair list."""

    def __init__(self):
        self.lib = _lib.load()
        h = vp()
        check(self.lib.ppf_contact_create(ctypes.byref(h)), "ppf_contact_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.lib.ppf_contact_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def assemble(self, t, x0, mass, fric, pairs, friction_mode=0, friction_eps=1e-5,
                 elastic=None, static=None):
        """t: device contact mesh (x = Newton iterate); x0, mass: device
        tensors; fric: dict v/e/f -> device tensors; pairs: (k, 4) int32 device
        tensor; elastic: (row_offset, col, val) device tensors or None;
        static: device collision mesh (with v/e/f_fric) for kind 4-6 pairs.
        Returns (grad, row_offset, rows, cols, vals) as device tensors (the
        block arrays are copies)."""
        from .ccd import static_struct
        s = mesh_struct(t)
        el = elastic or (None, None, None)
        ss = static_struct(static) if static is not None else None
        p = ContactParamsC(x0=_ptr(x0), mass=_ptr(mass), v_fric=_ptr(fric["v"]),
                           e_fric=_ptr(fric["e"]), f_fric=_ptr(fric["f"]),
                           friction_mode=friction_mode, friction_eps=friction_eps,
                           el_rp=_ptr(el[0]), el_col=_ptr(el[1]), el_val=_ptr(el[2]),
                           sm=ctypes.addressof(ss) if ss is not None else None)
        grad = torch.empty(3 * t["nv"], dtype=torch.float32, device=t["x"].device)
        nnz = ctypes.c_uint64()
        check(self.lib.ppf_contact_assemble(self._h, ctypes.byref(s), ctypes.byref(p),
                                            vp(pairs.data_ptr()), pairs.shape[0],
                                            vp(grad.data_ptr()), ctypes.byref(nnz),
                                            _stream()), "ppf_contact_assemble")
        rp, rows, cols, vals = vp(), vp(), vp(), vp()
        check(self.lib.ppf_contact_blocks(self._h, ctypes.byref(rp), ctypes.byref(rows),
                                          ctypes.byref(cols), ctypes.byref(vals),
                                          ctypes.byref(nnz)), "ppf_contact_blocks")
        n = nnz.value
        dev = t["x"].device
        out_rp = _copy_dev(rp.value, t["nv"] + 1, torch.int32, dev)
        out_rows = _copy_dev(rows.value, n, torch.int32, dev)
        out_cols = _copy_dev(cols.value, n, torch.int32, dev)
        out_vals = _copy_dev(vals.value, 9 * n, torch.float32, dev)
        return grad, out_rp, out_rows, out_cols, out_vals


def _copy_dev(ptr, count, dtype, dev):
    """Copy `count` elements from a library-owned device pointer."""
    out = torch.empty(max(1, count), dtype=dtype, device=dev)
    if count and ptr:
        nbytes = count * out.element_size()
        rc = _hip().hipMemcpy(vp(out.data_ptr()), vp(ptr), ctypes.c_size_t(nbytes), 3)
        if rc != 0:
            raise _lib.PPFError(f"hipMemcpy failed ({rc})")
    return out[:count]


_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        _HIP = ctypes.CDLL("libamdhip64.so")
        _HIP.hipMemcpy.restype = ctypes.c_int
        _HIP.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    return _HIP
