"""ctypes binding of the CPU oracle (liboracle.so) and of the reference-header
build (_ref/libppfref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The product never uses it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libppfref.so")

u32 = ctypes.c_uint32
f32 = ctypes.c_float
vp = ctypes.c_void_p


class OracleResult(ctypes.Structure):
    _fields_ = [("iter", u32), ("resid", f32), ("status", ctypes.c_int32),
                ("break_cause", ctypes.c_int32), ("break_iter", ctypes.c_int32),
                ("err0", f32)]


_lib = None
_ref = None


def load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(LIB)
        lib.oracle_symm3x3.argtypes = [vp, vp, vp]
        lib.oracle_bj_invert.argtypes = [u32, vp, vp]
        lib.oracle_bj_invert.restype = u32
        lib.oracle_spmv.argtypes = [u32, vp, vp, vp, vp, vp, vp]
        lib.oracle_merged_diag.argtypes = [u32, vp, vp, vp, vp, vp]
        lib.oracle_pcg.argtypes = [u32, vp, vp, vp, vp, vp, vp, f32, u32, vp,
                                   ctypes.POINTER(OracleResult)]
        lib.oracle_pcg.restype = ctypes.c_int
        lib.oracle_set_threads.argtypes = [ctypes.c_int]
        lib.oracle_max_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def load_ref():
    """The reference's own headers compiled host-side, or None if not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        lib = ctypes.CDLL(REF_LIB)
        lib.ref_symm3x3.argtypes = [vp, vp, vp]
        _ref = lib
    return _ref


def _p(a):
    return a.ctypes.data_as(vp)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def set_threads(n):
    load().oracle_set_threads(int(n))


def symm3x3(a9, lib=None):
    a9 = _f32(a9).reshape(-1, 9)
    val = np.zeros((len(a9), 3), np.float32)
    vec = np.zeros((len(a9), 9), np.float32)
    fn = (lib or load()).oracle_symm3x3 if lib is None else lib.ref_symm3x3
    for k in range(len(a9)):
        fn(_p(a9[k]), _p(val[k]), _p(vec[k]))
    return val, vec


def ref_symm3x3(a9):
    ref = load_ref()
    if ref is None:
        return None
    return symm3x3(a9, lib=ref)


def bj_invert(a9):
    a9 = _f32(a9).reshape(-1, 9)
    out = np.zeros_like(a9)
    bad = load().oracle_bj_invert(len(a9), _p(a9), _p(out))
    return out, int(bad)


def spmv(sys, x):
    x = _f32(x)
    y = np.zeros_like(x)
    ro, ci = _u32(sys["row_offset"]), _u32(sys["col_index"])
    val, cd = _f32(sys["value"]), _f32(sys["cdiag"])
    load().oracle_spmv(sys["n"], _p(ro), _p(ci), _p(val), _p(cd), _p(x), _p(y))
    return y


def pcg(sys, x0=None, tol=1e-3, max_iter=10000, hist=False):
    n = sys["n"]
    x = np.zeros(3 * n, np.float32) if x0 is None else _f32(x0).copy()
    ro, ci = _u32(sys["row_offset"]), _u32(sys["col_index"])
    val, cd, b = _f32(sys["value"]), _f32(sys["cdiag"]), _f32(sys["b"])
    h = np.zeros(max_iter, np.float32) if hist else None
    res = OracleResult()
    load().oracle_pcg(n, _p(ro), _p(ci), _p(val), _p(cd), _p(b), _p(x), tol,
                      max_iter, _p(h) if hist else None, ctypes.byref(res))
    out = dict(iter=res.iter, resid=res.resid, status=res.status,
               break_cause=res.break_cause, break_iter=res.break_iter,
               err0=res.err0, x=x)
    if hist:
        out["err_hist"] = h[:res.iter]
    return out


# ---- narrow phase -----------------------------------------------------------

class ContactMesh(ctypes.Structure):
    _fields_ = [("nv", u32), ("ne", u32), ("nf", u32), ("x", vp),
                ("edge", vp), ("face", vp),
                ("v2e_off", vp), ("v2e", vp), ("v2f_off", vp), ("v2f", vp),
                ("e2f_off", vp), ("e2f", vp),
                ("v_ghat", vp), ("v_off", vp), ("e_ghat", vp), ("e_off", vp),
                ("f_ghat", vp), ("f_off", vp),
                ("vflag", vp), ("vbody", vp), ("eflag", vp), ("fflag", vp)]


_FLOAT_FIELDS = {"x", "v_ghat", "v_off", "e_ghat", "e_off", "f_ghat", "f_off"}

PAIR_DTYPE = np.dtype([("kind", np.uint32), ("a", np.uint32), ("b", np.uint32),
                       ("count", np.uint32)])


def _mesh_struct(m):
    keep = {}
    fields = {}
    for name, _ in ContactMesh._fields_:
        if name in ("nv", "ne", "nf"):
            fields[name] = int(m[name])
            continue
        dt = np.float32 if name in _FLOAT_FIELDS else np.uint32
        a = np.ascontiguousarray(m[name], dtype=dt)
        keep[name] = a
        fields[name] = a.ctypes.data_as(vp)
    return ContactMesh(**fields), keep


def _load_narrow():
    lib = load()
    if not hasattr(lib, "_narrow_ready"):
        lib.oracle_proximity.argtypes = [ctypes.POINTER(ContactMesh), vp,
                                         ctypes.c_uint64]
        lib.oracle_proximity.restype = ctypes.c_uint64
        lib.oracle_bvh_build.argtypes = [ctypes.POINTER(ContactMesh), ctypes.c_int,
                                         vp, vp, ctypes.POINTER(u32)]
        for fn in ("oracle_pe_coeff", "oracle_pt_coeff", "oracle_ee_coeff"):
            getattr(lib, fn).restype = None
        lib._narrow_ready = True
    return lib


def proximity(m):
    """Stage-0 proximity pairs (structured array kind, a, b, count)."""
    lib = _load_narrow()
    s, keep = _mesh_struct(m)
    cap = 1 << 16
    while True:
        out = np.zeros(cap, PAIR_DTYPE)
        n = lib.oracle_proximity(ctypes.byref(s), out.ctypes.data_as(vp), cap)
        if n <= cap:
            return out[:n]
        cap = int(n)


def bvh_build(m, kind):
    """(nodes (2n-1, 2) uint32, boxes (2n-1, 6) float32, root)."""
    lib = _load_narrow()
    s, keep = _mesh_struct(m)
    n = [m["nf"], m["ne"], m["nv"]][kind]
    nn = max(1, 2 * n - 1)
    nodes = np.zeros((nn, 2), np.uint32)
    boxes = np.zeros((nn, 6), np.float32)
    root = u32()
    lib.oracle_bvh_build(ctypes.byref(s), kind, nodes.ctypes.data_as(vp),
                         boxes.ctypes.data_as(vp), ctypes.byref(root))
    return nodes, boxes, root.value


def coeff(kind, pts, lib_name="oracle"):
    """Distance coefficients for a batch: kind 'pe' (p,e0,e1), 'pt' (p,t0,t1,t2),
    'ee' (a0,a1,b0,b1); pts shape (k, m, 3). lib_name 'ref' uses the
    reference-header build."""
    lib = _load_narrow() if lib_name == "oracle" else load_ref()
    fn = getattr(lib, f"{lib_name}_{kind}_coeff")
    fn.restype = None
    pts = _f32(pts)
    k, mm = pts.shape[0], pts.shape[1]
    nout = {"pe": 2, "pt": 3, "ee": 4}[kind]
    out = np.zeros((k, nout), np.float32)
    for i in range(k):
        args = [pts[i, j].ctypes.data_as(vp) for j in range(mm)]
        fn(*args, out[i].ctypes.data_as(vp))
    return out


# ---- contact assembly -------------------------------------------------------

class ContactParams(ctypes.Structure):
    _fields_ = [("x0", vp), ("mass", vp), ("v_fric", vp), ("e_fric", vp),
                ("f_fric", vp), ("friction_mode", ctypes.c_int),
                ("friction_eps", ctypes.c_float), ("el_rp", vp), ("el_col", vp),
                ("el_val", vp), ("sm", vp)]


def contact_assemble(m, x0, mass, fric, pairs, friction_mode=0, friction_eps=1e-5,
                     elastic=None, static=None):
    """(grad (3nv,), rows, cols, vals (nnz, 9)) in the oracle's order."""
    lib = load()
    if not hasattr(lib, "_asm_ready"):
        lib.oracle_contact_assemble.argtypes = [ctypes.POINTER(ContactMesh),
                                                ctypes.POINTER(ContactParams), vp,
                                                ctypes.c_uint64, vp, vp, vp, vp,
                                                ctypes.c_uint64]
        lib.oracle_contact_assemble.restype = ctypes.c_uint64
        lib._asm_ready = True
    s, keep = _mesh_struct(m)
    x0 = _f32(x0).reshape(-1)
    mass = _f32(mass)
    fv, fe, ff = _f32(fric["v"]), _f32(fric["e"]), _f32(fric["f"])
    el = elastic or (None, None, None)
    el_rp = _u32(el[0]) if el[0] is not None else None
    el_col = _u32(el[1]) if el[1] is not None else None
    el_val = _f32(el[2]).reshape(-1) if el[2] is not None else None
    ss, keep2 = static_mesh_struct(static)
    p = ContactParams(x0=_p(x0), mass=_p(mass), v_fric=_p(fv), e_fric=_p(fe),
                      f_fric=_p(ff), friction_mode=friction_mode,
                      friction_eps=friction_eps,
                      el_rp=_p(el_rp) if el_rp is not None else None,
                      el_col=_p(el_col) if el_col is not None else None,
                      el_val=_p(el_val) if el_val is not None else None,
                      sm=ctypes.addressof(ss) if static is not None else None)
    pr = np.ascontiguousarray(pairs).view(np.uint32).reshape(-1, 4).copy()
    grad = np.zeros(3 * int(m["nv"]), np.float32)
    cap = 16 * max(1, len(pr))
    rows = np.zeros(cap, np.uint32)
    cols = np.zeros(cap, np.uint32)
    vals = np.zeros((cap, 9), np.float32)
    n = lib.oracle_contact_assemble(ctypes.byref(s), ctypes.byref(p), _p(pr), len(pr),
                                    _p(grad), _p(rows), _p(cols), _p(vals), cap)
    return grad, rows[:n], cols[:n], vals[:n]


def analytic_assemble(m, x0, mass, v_fric, spheres=(), floors=(), friction_mode=0,
                      friction_eps=1e-5, constraint_tol=0.01, elastic=None, grad=None,
                      cdiag=None):
    """Analytic sphere / floor terms of every vertex (contact.cu:1154-1382)
    added to grad (3 nv) and cdiag (nv, 9), both zero when not given.
    Returns (grad, cdiag, contacts)."""
    lib = load()
    if not hasattr(lib, "_ana_ready"):
        lib.oracle_analytic_assemble.argtypes = [ctypes.POINTER(ContactMesh),
                                                 ctypes.POINTER(ContactParams), vp, u32, vp,
                                                 u32, f32, vp, vp]
        lib.oracle_analytic_assemble.restype = u32
        lib._ana_ready = True
    s, keep = _mesh_struct(m)
    x0 = _f32(x0).reshape(-1)
    mass = _f32(mass)
    fv = _f32(v_fric)
    el = elastic or (None, None, None)
    el_rp = _u32(el[0]) if el[0] is not None else None
    el_col = _u32(el[1]) if el[1] is not None else None
    el_val = _f32(el[2]).reshape(-1) if el[2] is not None else None
    p = ContactParams(x0=_p(x0), mass=_p(mass), v_fric=_p(fv), e_fric=None, f_fric=None,
                      friction_mode=friction_mode, friction_eps=friction_eps,
                      el_rp=_p(el_rp) if el_rp is not None else None,
                      el_col=_p(el_col) if el_col is not None else None,
                      el_val=_p(el_val) if el_val is not None else None, sm=None)
    sp, nsp, fl, nfl = collider_arrays({"spheres": list(spheres), "floors": list(floors)})
    nv = int(m["nv"])
    grad = np.zeros(3 * nv, np.float32) if grad is None else _f32(grad).reshape(-1).copy()
    cdiag = np.zeros((nv, 9), np.float32) if cdiag is None else _f32(cdiag).reshape(nv, 9).copy()
    n = lib.oracle_analytic_assemble(ctypes.byref(s), ctypes.byref(p), ctypes.cast(sp, vp), nsp,
                                     ctypes.cast(fl, vp), nfl, constraint_tol, _p(grad),
                                     _p(cdiag))
    return grad, cdiag, int(n)


def analytic_toi(x0, x1, vflag, vmass, spheres=(), floors=(), max_t=1.25):
    """The analytic part of the line search (contact.cu:2318-2470): (toi in
    [0, max_t] over the swept step x0 -> x0 + max_t (x1 - x0), the smallest
    pinned vertex driven into a static collider or None)."""
    lib = load()
    if not hasattr(lib, "_anatoi_ready"):
        lib.oracle_analytic_toi.argtypes = [u32, vp, vp, vp, vp, vp, u32, vp, u32, f32, vp, vp]
        lib.oracle_analytic_toi.restype = None
        lib._anatoi_ready = True
    x0 = _f32(x0).reshape(-1)
    x1 = _f32(x1).reshape(-1)
    vf = _u32(vflag)
    vm = _f32(vmass)
    sp, nsp, fl, nfl = collider_arrays({"spheres": list(spheres), "floors": list(floors)})
    toi = np.array([max_t], np.float32)
    bad = np.array([0xFFFFFFFF], np.uint32)
    lib.oracle_analytic_toi(len(vm), _p(x0), _p(x1), _p(vf), _p(vm), ctypes.cast(sp, vp), nsp,
                            ctypes.cast(fl, vp), nfl, max_t, _p(toi), _p(bad))
    return float(toi[0]), (None if bad[0] == 0xFFFFFFFF else int(bad[0]))


# ---- CCD / line search ------------------------------------------------------

class StaticMesh(ctypes.Structure):
    _fields_ = [("nv", u32), ("ne", u32), ("nf", u32), ("x", vp), ("edge", vp),
                ("face", vp), ("v_ghat", vp), ("v_off", vp), ("e_ghat", vp),
                ("e_off", vp), ("f_ghat", vp), ("f_off", vp), ("v_fric", vp),
                ("e_fric", vp), ("f_fric", vp)]


class LsParams(ctypes.Structure):
    _fields_ = [("x1", vp), ("vmass", vp), ("emass", vp), ("fmass", vp),
                ("max_t", f32), ("ccd_eps", f32)]


def _load_ccd():
    lib = load()
    if not hasattr(lib, "_ccd_ready"):
        lib.oracle_ccd.argtypes = [ctypes.c_int, vp, vp, f32, f32, f32, f32,
                                   ctypes.POINTER(ctypes.c_int)]
        lib.oracle_ccd.restype = f32
        lib.oracle_line_search.argtypes = [ctypes.POINTER(ContactMesh),
                                           ctypes.POINTER(StaticMesh),
                                           ctypes.POINTER(LsParams),
                                           ctypes.POINTER(ctypes.c_int)]
        lib.oracle_line_search.restype = f32
        lib._ccd_ready = True
    return lib


def ccd(kind, x0, x1, offset, ghat, max_t=1.25, ccd_eps=1e-7):
    """Single-pair ACCD: kind 0 point-triangle, 1 point-edge, 2 point-point,
    3 edge-edge; x0/x1 (columns, 3). Returns (toi, overlap)."""
    lib = _load_ccd()
    a, b = _f32(x0).reshape(-1), _f32(x1).reshape(-1)
    ov = ctypes.c_int()
    t = lib.oracle_ccd(kind, _p(a), _p(b), offset, ghat, max_t, ccd_eps,
                       ctypes.byref(ov))
    return float(t), bool(ov.value)


def static_mesh_struct(sm):
    keep = {}
    if sm is None:
        return StaticMesh(nv=0, ne=0, nf=0), keep
    fields = {"nv": int(sm["nv"]), "ne": int(sm["ne"]), "nf": int(sm["nf"])}
    for name in ("x", "edge", "face", "v_ghat", "v_off", "e_ghat", "e_off",
                 "f_ghat", "f_off", "v_fric", "e_fric", "f_fric"):
        if name not in sm:
            continue
        dt = np.uint32 if name in ("edge", "face") else np.float32
        a = np.ascontiguousarray(sm[name], dtype=dt)
        if a.size == 0:
            a = np.zeros(1, dt)
        keep[name] = a
        fields[name] = _p(a)
    return StaticMesh(**fields), keep


def line_search(m, x1, vmass, emass, fmass, static=None, max_t=1.25, ccd_eps=1e-7):
    """contact::line_search over the mesh from m['x'] to x1: (fraction, overlap)."""
    lib = _load_ccd()
    s, keep = _mesh_struct(m)
    ss, keep2 = static_mesh_struct(static)
    x1 = _f32(x1).reshape(-1)
    vm, em, fm = _f32(vmass), _f32(emass), _f32(fmass)
    p = LsParams(x1=_p(x1), vmass=_p(vm), emass=_p(em), fmass=_p(fm), max_t=max_t,
                 ccd_eps=ccd_eps)
    ov = ctypes.c_int()
    t = lib.oracle_line_search(ctypes.byref(s), ctypes.byref(ss), ctypes.byref(p),
                               ctypes.byref(ov))
    return float(t), bool(ov.value)


# ---- intersection check -----------------------------------------------------

ISECT_DTYPE = np.dtype([("type", np.uint32), ("elem0", np.uint32), ("elem1", np.uint32),
                        ("num_verts0", np.uint32), ("num_verts1", np.uint32),
                        ("positions", np.float32, (15,))])


def _load_isect():
    lib = load()
    if not hasattr(lib, "_isect_ready"):
        lib.oracle_intersection_check.argtypes = [ctypes.POINTER(ContactMesh),
                                                  ctypes.POINTER(StaticMesh), vp, vp, vp,
                                                  u32, vp, u32]
        lib.oracle_intersection_check.restype = u32
        lib._isect_ready = True
    return lib


def intersection_check(m, vmass, emass, fmass, static=None, surface=None, cap=256):
    """contact::check_intersection at m['x']: (count, records[:min(count, cap)])."""
    lib = _load_isect()
    s, keep = _mesh_struct(m)
    ss, keep2 = static_mesh_struct(static)
    vm, em, fm = _f32(vmass), _f32(emass), _f32(fmass)
    rec = np.zeros(max(1, cap), ISECT_DTYPE)
    nsurf = int(m["nv"]) if surface is None else int(surface)
    n = lib.oracle_intersection_check(ctypes.byref(s), ctypes.byref(ss), _p(vm), _p(em),
                                      _p(fm), nsurf, rec.ctypes.data_as(vp), cap)
    return int(n), rec[:min(int(n), cap)]


# ---- elastic / inertia assembly ---------------------------------------------

class Shell(ctypes.Structure):
    _fields_ = [("nv", u32), ("nf", u32), ("nh", u32), ("x", vp), ("x0", vp),
                ("target", vp), ("face", vp), ("hinge", vp), ("v2f_off", vp), ("v2f", vp),
                ("vmass", vp), ("varea", vp), ("vflag", vp), ("fmass", vp),
                ("inv_rest", vp), ("f_mu", vp), ("f_lambda", vp), ("fflag", vp),
                ("h_rest", vp), ("h_length", vp), ("h_area", vp), ("h_bend", vp),
                ("hflag", vp), ("f_strainlimit", vp), ("f_shrink", vp),
                ("f_plasticity", vp), ("f_plastic_threshold", vp), ("h_plasticity", vp),
                ("h_plastic_threshold", vp)]


class ElasticParams(ctypes.Structure):
    _fields_ = [("dt", f32), ("air_density", f32), ("air_friction", f32),
                ("wind", f32 * 3)]


_SHELL_U32 = {"face", "hinge", "v2f_off", "v2f", "vflag", "fflag", "hflag"}
# per-face / per-hinge arrays that may be absent (null): strain limits, rest
# shrink, plasticity
_SHELL_OPT = {"f_strainlimit", "f_shrink", "f_plasticity", "f_plastic_threshold",
              "h_plasticity", "h_plastic_threshold"}


def shell_struct(sh):
    keep, fields = {}, {}
    for name, _ in Shell._fields_:
        if name in ("nv", "nf", "nh"):
            fields[name] = int(sh[name])
            continue
        if name in _SHELL_OPT and sh.get(name) is None:
            fields[name] = None
            continue
        dt = np.uint32 if name in _SHELL_U32 else np.float32
        a = np.ascontiguousarray(sh[name], dtype=dt).reshape(-1)
        if a.size == 0:
            a = np.zeros(1, dt)
        keep[name] = a
        fields[name] = a.ctypes.data_as(vp)
    return Shell(**fields), keep


def _load_elastic():
    lib = load()
    if not hasattr(lib, "_elastic_ready"):
        lib.oracle_elastic_local.argtypes = [ctypes.c_int, vp, f32, vp, vp]
        lib.oracle_elastic_local.restype = None
        lib.oracle_dihedral_angle.argtypes = [vp]
        lib.oracle_dihedral_angle.restype = f32
        lib.oracle_elastic_assemble.argtypes = [ctypes.POINTER(Shell),
                                                ctypes.POINTER(ElasticParams), vp, vp, vp,
                                                vp, vp]
        lib.oracle_elastic_assemble.restype = u32
        lib._elastic_ready = True
    return lib


def elastic_local(kind, inp, s0):
    """kind 0 stretch / 1 shear (F (3,2)) -> (grad (3,2), hess (6,6)); kind 2
    hinge (4 remapped points (4,3), s0 = rest angle) -> (force (4,3), hess (12,12))."""
    lib = _load_elastic()
    if kind in (0, 1):
        a = np.asfortranarray(np.asarray(inp, np.float32).reshape(3, 2)).ravel(order="F")
        g = np.zeros(6, np.float32)
        h = np.zeros(36, np.float32)
        lib.oracle_elastic_local(kind, _p(a), s0, _p(g), _p(h))
        return g.reshape(2, 3).T, h.reshape(6, 6).T
    a = _f32(inp).reshape(-1)
    g = np.zeros(12, np.float32)
    h = np.zeros(144, np.float32)
    lib.oracle_elastic_local(kind, _p(a), s0, _p(g), _p(h))
    return g.reshape(4, 3), h.reshape(12, 12)


def ref_elastic_local(kind, inp, s0):
    """The reference's own BaraffWitkin / face_compute_force_hessian (built
    from the reference headers into oracle/_ref); kind 2 takes the hinge's four
    points in STORED order and returns the force in the remapped order."""
    lib = load_ref()
    if not hasattr(lib, "_elastic_ready"):
        lib.ref_bw.argtypes = [ctypes.c_int, vp, f32, vp, vp]
        lib.ref_bw.restype = None
        lib.ref_hinge.argtypes = [vp, f32, vp, vp]
        lib.ref_hinge.restype = None
        lib._elastic_ready = True
    if kind in (0, 1):
        a = np.asarray(inp, np.float32).reshape(3, 2).ravel(order="F").copy()
        g = np.zeros(6, np.float32)
        h = np.zeros(36, np.float32)
        lib.ref_bw(kind, _p(a), s0, _p(g), _p(h))
        return g.reshape(2, 3).T, h.reshape(6, 6).T
    a = _f32(inp).reshape(-1).copy()
    g = np.zeros(12, np.float32)
    h = np.zeros(144, np.float32)
    lib.ref_hinge(_p(a), s0, _p(g), _p(h))
    return g.reshape(4, 3), h.reshape(12, 12)


def dihedral_angle(pts):
    lib = _load_elastic()
    a = _f32(pts).reshape(-1)
    return float(lib.oracle_dihedral_angle(_p(a)))


def elastic_assemble(sh, rp, ci, dt, air_density=1e-3, air_friction=0.2, wind=(0, 0, 0)):
    """Elastic + inertia assembly over the upper pattern (rp, ci):
    (force (nv,3), cdiag (nv,3,3) column-major blocks, val (nnz, 9), missing)."""
    lib = _load_elastic()
    s, keep = shell_struct(sh)
    p = ElasticParams(dt=dt, air_density=air_density, air_friction=air_friction,
                      wind=(f32 * 3)(*wind))
    rp = _u32(rp)
    ci = _u32(ci)
    nv = int(sh["nv"])
    force = np.zeros(3 * nv, np.float32)
    cdiag = np.zeros(9 * nv, np.float32)
    val = np.zeros(9 * max(1, len(ci)), np.float32)
    miss = lib.oracle_elastic_assemble(ctypes.byref(s), ctypes.byref(p), _p(rp), _p(ci),
                                       _p(force), _p(cdiag), _p(val))
    return force, cdiag, val[:9 * len(ci)].reshape(-1, 9), int(miss)


def _load_strainlimit():
    lib = load()
    if not hasattr(lib, "_sl_ready"):
        lib.oracle_svd32.argtypes = [u32, vp, vp, vp, vp]
        lib.oracle_svd32.restype = None
        lib.oracle_strainlimit_faces.argtypes = [u32] + [vp] * 11 + [f32] + [vp] * 5
        lib.oracle_strainlimit_faces.restype = None
        lib.oracle_strainlimit_assemble.argtypes = [u32, u32] + [vp] * 10 + [f32, vp, vp]
        lib.oracle_strainlimit_assemble.restype = None
        lib.oracle_strainlimit_toi.argtypes = [u32] + [vp] * 7 + [f32]
        lib.oracle_strainlimit_toi.restype = f32
        lib._sl_ready = True
    return lib


def svd32(F):
    """One-rotation Jacobi SVD of 3x2 matrices F (n, 3, 2):
    U (n, 3, 2), S (n, 2) ascending, V (n, 2, 2)."""
    lib = _load_strainlimit()
    F = np.asarray(F, np.float32).reshape(-1, 3, 2)
    n = F.shape[0]
    Fc = np.ascontiguousarray(F.transpose(0, 2, 1)).reshape(-1)  # column-major
    U = np.zeros(6 * n, np.float32)
    S = np.zeros(2 * n, np.float32)
    V = np.zeros(4 * n, np.float32)
    lib.oracle_svd32(n, _p(Fc), _p(U), _p(S), _p(V))
    return (U.reshape(n, 2, 3).transpose(0, 2, 1), S.reshape(n, 2),
            V.reshape(n, 2, 2).transpose(0, 2, 1))


def _sl_arrays(sh):
    return (_u32(sh["face"]), _f32(sh["inv_rest"]), _u32(sh["fflag"]), _f32(sh["fmass"]),
            _f32(sh["f_strainlimit"]),
            _f32(sh["f_shrink"]) if sh.get("f_shrink") is not None else None)


def strainlimit_faces(sh, x, x0, rp, ci, el, eig_eps=1e-2):
    """Per-face strain-limit terms at positions x (stiffness from x0 and the
    elastic blocks el over the upper pattern rp / ci): forces (nf, 9), upper
    blocks (nf, 6, 9), stiffness (nf,), singular values (nf, 2), active (nf,)."""
    lib = _load_strainlimit()
    nf = int(sh["nf"])
    face, inv, ffl, fm, sl, shr = _sl_arrays(sh)
    x, x0, rp, ci, el = _f32(x), _f32(x0), _u32(rp), _u32(ci), _f32(el)
    fo = np.zeros(9 * nf, np.float32)
    bo = np.zeros(54 * nf, np.float32)
    st = np.zeros(nf, np.float32)
    S = np.zeros(2 * nf, np.float32)
    act = np.zeros(nf, np.uint8)
    lib.oracle_strainlimit_faces(nf, _p(face), _p(inv), _p(ffl), _p(fm), _p(sl),
                                 _p(shr) if shr is not None else None, _p(x), _p(x0), _p(rp),
                                 _p(ci), _p(el), eig_eps, _p(fo), _p(bo), _p(st), _p(S),
                                 _p(act))
    return fo.reshape(nf, 9), bo.reshape(nf, 6, 9), st, S.reshape(nf, 2), act


def strainlimit_assemble(sh, x, x0, rp, ci, force, value, eig_eps=1e-2):
    """Adds the strain-limit terms to copies of force (3 nv) and value (nnz, 9)."""
    lib = _load_strainlimit()
    face, inv, ffl, fm, sl, shr = _sl_arrays(sh)
    force = np.array(force, np.float32).reshape(-1)
    value = np.array(value, np.float32).reshape(-1, 9)
    x, x0, rp, ci = _f32(x), _f32(x0), _u32(rp), _u32(ci)
    lib.oracle_strainlimit_assemble(int(sh["nv"]), int(sh["nf"]), _p(face), _p(inv), _p(ffl),
                                    _p(fm), _p(sl), _p(shr) if shr is not None else None,
                                    _p(x), _p(x0), _p(rp), _p(ci), eig_eps, _p(force),
                                    _p(value))
    return force, value


def strainlimit_toi(sh, x0, x1, max_t=1.25):
    """Largest t <= max_t keeping every limited face below its limit along
    x0 + t (x1 - x0) (unnormalized)."""
    lib = _load_strainlimit()
    face, inv, ffl, _, sl, shr = _sl_arrays(sh)
    x0, x1 = _f32(x0), _f32(x1)
    return float(lib.oracle_strainlimit_toi(int(sh["nf"]), _p(face), _p(inv), _p(ffl), _p(sl),
                                            _p(shr) if shr is not None else None, _p(x0),
                                            _p(x1), max_t))


def plasticity(sh, x, dt):
    """Rest-shape creep at positions x: (inv_rest (nf, 4), h_rest (nh,)) after
    one committed step (copies; sh is not modified)."""
    lib = load()
    if not hasattr(lib, "_plastic_ready"):
        lib.oracle_plasticity.argtypes = [u32, u32] + [vp] * 8 + [f32, vp, vp]
        lib.oracle_plasticity.restype = None
        lib._plastic_ready = True
    inv = np.array(sh["inv_rest"], np.float32).reshape(-1, 4)
    hr = np.array(sh["h_rest"], np.float32).reshape(-1)
    opt = {k: (_f32(sh[k]) if sh.get(k) is not None else None)
           for k in ("f_plasticity", "f_plastic_threshold", "h_plasticity",
                     "h_plastic_threshold")}
    face, hinge, hflag, x = _u32(sh["face"]), _u32(sh["hinge"]), _u32(sh["hflag"]), _f32(x)
    lib.oracle_plasticity(int(sh["nf"]), int(sh["nh"]), _p(face), _p(hinge) if hinge.size else None,
                          _p(hflag) if hflag.size else None, _p(x),
                          *(_p(opt[k]) if opt[k] is not None else None
                            for k in ("f_plasticity", "f_plastic_threshold", "h_plasticity",
                                      "h_plastic_threshold")),
                          dt, _p(inv), _p(hr) if hr.size else None)
    return inv, hr


def static_proximity(m, sm, vmass, emass, fmass):
    """Collision-mesh proximity pairs (kinds 4-6), queries in index order."""
    lib = _load_narrow()
    if not hasattr(lib, "_static_ready"):
        lib.oracle_static_proximity.argtypes = [ctypes.POINTER(ContactMesh),
                                                ctypes.POINTER(StaticMesh), vp, vp, vp, vp,
                                                ctypes.c_uint64]
        lib.oracle_static_proximity.restype = ctypes.c_uint64
        lib._static_ready = True
    s, keep = _mesh_struct(m)
    ss, keep2 = static_mesh_struct(sm)
    vm, em, fm = _f32(vmass), _f32(emass), _f32(fmass)
    cap = 1 << 14
    while True:
        out = np.zeros(cap, PAIR_DTYPE)
        n = lib.oracle_static_proximity(ctypes.byref(s), ctypes.byref(ss), _p(vm), _p(em),
                                        _p(fm), out.ctypes.data_as(vp), cap)
        if n <= cap:
            return out[:n]
        cap = int(n)


# ---- one time step ------------------------------------------------------------

class OracleSphere(ctypes.Structure):
    """Analytic sphere (data.hpp:449-458)."""
    _fields_ = [("center", f32 * 3), ("ghat", f32), ("friction", f32), ("radius", f32),
                ("thickness", f32), ("bowl", ctypes.c_bool), ("reverse", ctypes.c_bool),
                ("kinematic", ctypes.c_bool)]


class OracleFloor(ctypes.Structure):
    """Analytic floor (data.hpp:460-467)."""
    _fields_ = [("ground", f32 * 3), ("ghat", f32), ("friction", f32), ("thickness", f32),
                ("up", f32 * 3), ("kinematic", ctypes.c_bool)]


def collider_arrays(sc, sphere_t=OracleSphere, floor_t=OracleFloor):
    """ctypes arrays of a scene's analytic colliders (sc['spheres'] /
    sc['floors']: lists of dicts with the Sphere / Floor fields)."""
    sp = sc.get("spheres") or []
    fl = sc.get("floors") or []
    a = (sphere_t * max(1, len(sp)))()
    for k, d in enumerate(sp):
        a[k] = sphere_t(center=(f32 * 3)(*d["center"]), ghat=d["ghat"], friction=d["friction"],
                        radius=d["radius"], thickness=d.get("thickness", 0.0),
                        bowl=bool(d.get("bowl", False)), reverse=bool(d.get("reverse", False)),
                        kinematic=bool(d.get("kinematic", False)))
    b = (floor_t * max(1, len(fl)))()
    for k, d in enumerate(fl):
        b[k] = floor_t(ground=(f32 * 3)(*d["ground"]), ghat=d["ghat"], friction=d["friction"],
                       thickness=d.get("thickness", 0.0), up=(f32 * 3)(*d.get("up", (0, 1, 0))),
                       kinematic=bool(d.get("kinematic", False)))
    return a, len(sp), b, len(fl)


class OracleScene(ctypes.Structure):
    _fields_ = [("mesh", ContactMesh), ("shell", Shell), ("collision", StaticMesh),
                ("prev", vp), ("emass", vp), ("v_fric", vp), ("e_fric", vp), ("f_fric", vp),
                ("fix_pos", vp), ("fix_delta", vp), ("spheres", vp), ("nsphere", u32),
                ("floors", vp), ("nfloor", u32)]


class OracleStepParams(ctypes.Structure):
    _fields_ = [("dt", f32), ("prev_dt", f32), ("gravity", f32 * 3), ("cg_tol", f32),
                ("cg_max_iter", u32), ("target_toi", f32), ("line_search_max_t", f32),
                ("ccd_eps", f32), ("max_dx", f32), ("friction_eps", f32),
                ("friction_mode", ctypes.c_int32), ("min_newton_steps", u32),
                ("max_newton_steps", u32), ("air_density", f32), ("air_friction", f32),
                ("wind", f32 * 3), ("disable_contact", ctypes.c_int32),
                ("check_intersection", ctypes.c_int32), ("constraint_tol", f32),
                ("eiganalysis_eps", f32)]


class OracleStepResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("newton_steps", u32), ("pcg_iter_total", u32),
                ("pcg_iter_max", u32), ("pairs_max", u32), ("intersections", u32),
                ("toi_advanced", f32), ("dt", f32)]


class OracleStepper:
    """The oracle's time stepping of a scene dict (scenes.drape_scene layout);
    scene['mesh']['x'] and scene['prev'] are advanced in place."""

    def __init__(self, sc):
        lib = load()
        if not hasattr(lib, "_step_ready"):
            lib.oracle_step.argtypes = [ctypes.POINTER(OracleScene),
                                        ctypes.POINTER(OracleStepParams),
                                        ctypes.POINTER(OracleStepResult)]
            lib.oracle_step.restype = ctypes.c_int
            lib._step_ready = True
        self.lib = lib
        self.sc = sc
        sc["mesh"]["x"] = np.ascontiguousarray(sc["mesh"]["x"], np.float32)
        sc["prev"] = np.ascontiguousarray(sc["prev"], np.float32)
        self.m, self.keep_m = _mesh_struct(sc["mesh"])
        # _mesh_struct copies x; point it at the live array instead
        self.m.x = _p(sc["mesh"]["x"])
        shell = dict(sc["shell"])
        for k in ("x", "x0", "target"):
            shell.setdefault(k, np.zeros(1, np.float32))
        self.s, self.keep_s = shell_struct(shell)
        self.st, self.keep_st = static_mesh_struct(sc.get("static"))
        self.keep = [_f32(sc["emass"]), _f32(sc["fric"]["v"]), _f32(sc["fric"]["e"]),
                     _f32(sc["fric"]["f"])]
        # fix_pos is rewound in place by the step: keep it in the scene dict
        if sc.get("fix_pos") is not None:
            sc["fix_pos"] = np.ascontiguousarray(sc["fix_pos"], np.float32)
        self.fix = sc["fix_pos"].reshape(-1) if sc.get("fix_pos") is not None else None
        self.delta = (_f32(sc["fix_delta"]).reshape(-1)
                      if sc.get("fix_delta") is not None else None)
        self.scene = OracleScene(mesh=self.m, shell=self.s, collision=self.st,
                                 prev=_p(sc["prev"]), emass=_p(self.keep[0]),
                                 v_fric=_p(self.keep[1]), e_fric=_p(self.keep[2]),
                                 f_fric=_p(self.keep[3]),
                                 fix_pos=_p(self.fix) if self.fix is not None else None,
                                 fix_delta=_p(self.delta) if self.delta is not None else None)

    def trace(self):
        """Per Newton iteration of the last advance: (PCG iterations, max_dx
        step scale, CCD toi, largest |dx|)."""
        lib = self.lib
        lib.oracle_step_trace.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        lib.oracle_step_trace.restype = ctypes.c_uint32
        out = np.zeros(4 * 64, np.float32)
        k = lib.oracle_step_trace(out.ctypes.data, 64)
        return out[:4 * min(k, 64)].reshape(-1, 4)

    def last_system(self):
        """The last Newton system the oracle solved: row_offset, col_index,
        value (nnz, 9), cdiag (n, 9), b (3n); eliminated blocks are zero."""
        lib = self.lib
        lib.oracle_last_system.argtypes = [ctypes.POINTER(u32), vp, vp, vp, vp, vp]
        lib.oracle_last_system.restype = u32
        n = u32()
        nnz = lib.oracle_last_system(ctypes.byref(n), None, None, None, None, None)
        n = n.value
        out = {"row_offset": np.zeros(n + 1, np.uint32), "col_index": np.zeros(nnz, np.uint32),
               "value": np.zeros((nnz, 9), np.float32), "cdiag": np.zeros((n, 9), np.float32),
               "b": np.zeros(3 * n, np.float32)}
        lib.oracle_last_system(None, *(_p(out[k]) for k in ("row_offset", "col_index", "value",
                                                             "cdiag", "b")))
        return out

    def advance(self, params):
        kw = {k: v for k, v in params.items() if k in dict(OracleStepParams._fields_)}
        kw["gravity"] = (f32 * 3)(*params["gravity"])
        kw["wind"] = (f32 * 3)(*params["wind"])
        p = OracleStepParams(**kw)
        r = OracleStepResult()
        # the analytic colliders as they are now (kinematic ones move)
        sp, nsp, fl, nfl = collider_arrays(self.sc)
        self.scene.spheres = ctypes.cast(sp, vp)
        self.scene.nsphere = nsp
        self.scene.floors = ctypes.cast(fl, vp)
        self.scene.nfloor = nfl
        self.lib.oracle_step(ctypes.byref(self.scene), ctypes.byref(p), ctypes.byref(r))
        return {k: getattr(r, k) for k, _ in OracleStepResult._fields_}
