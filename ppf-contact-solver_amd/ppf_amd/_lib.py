"""ctypes binding of libppf_mi355x.so (the C ABI declared in include/ppf_mi355x.h).

The library is the product: there is no fallback. If it is missing or was not
built for this machine, loading raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libppf_mi355x.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include",
                           "ppf_mi355x.h")

u32 = ctypes.c_uint32
i32 = ctypes.c_int32
f32 = ctypes.c_float
vp = ctypes.c_void_p


class PcgResult(ctypes.Structure):
    """ppf_pcg_result (include/ppf_mi355x.h)."""
    _fields_ = [("iter", u32), ("resid", f32), ("status", i32),
                ("break_cause", i32), ("break_iter", i32), ("err0", f32)]


class WalkStats(ctypes.Structure):
    """ppf_walk_stats (include/ppf_mi355x.h)."""
    _fields_ = [("launches", ctypes.c_uint64), ("queries", ctypes.c_uint64),
                ("nodes", ctypes.c_uint64), ("leaves", ctypes.c_uint64),
                ("sweeps", ctypes.c_uint64), ("ms", ctypes.c_double)]


WALK_KINDS = ("ls_vertex", "ls_edge", "prox_vertex", "prox_edge")

_PROTOS = {
    "ppf_last_error": (ctypes.c_char_p, []),
    "ppf_walk_stats_enable": (ctypes.c_int, [ctypes.c_int]),
    "ppf_walk_stats_read": (ctypes.c_int, [ctypes.POINTER(WalkStats)]),
    "ppf_alloc_counts": (None, [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "ppf_device_alloc": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_size_t]),
    "ppf_device_free": (None, [vp]),
    "ppf_device_info": (ctypes.c_int, [ctypes.c_char_p, u32, ctypes.POINTER(u32)]),
    "ppf_pcg_plan_create": (ctypes.c_int, [u32, u32, vp, vp, vp, ctypes.POINTER(vp)]),
    "ppf_pcg_plan_destroy": (None, [vp]),
    "ppf_pcg_plan_create_batch": (ctypes.c_int, [u32, u32, vp, vp, u32, vp, vp,
                                                 ctypes.POINTER(vp)]),
    "ppf_pcg_plan_set_tiles": (ctypes.c_int, [vp, u32, u32, vp]),
    "ppf_pcg_plan_set_order": (ctypes.c_int, [vp, vp, vp]),
    "ppf_stepper_solve_order": (ctypes.c_int, [vp, vp, vp]),
    "ppf_pcg_plan_autotune": (ctypes.c_int, [vp, vp, vp, vp, f32, vp, ctypes.POINTER(f32), vp]),
    "ppf_pcg_solve_batch": (ctypes.c_int, [vp, vp, vp, vp, vp, f32, u32, vp, vp]),
    "ppf_pcg_plan_repattern": (ctypes.c_int, [vp, u32, vp, vp, vp]),
    "ppf_lpc_weight_count": (u32, []),
    "ppf_lpc_create": (ctypes.c_int, [ctypes.POINTER(vp)]),
    "ppf_lpc_destroy": (None, [vp]),
    "ppf_lpc_set_weights": (ctypes.c_int, [vp, vp, u32, vp]),
    "ppf_lpc_get_weights": (ctypes.c_int, [vp, vp, u32, vp]),
    "ppf_lpc_factor": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
    "ppf_lpc_outputs": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
    "ppf_lpc_gradient": (ctypes.c_int, [vp, vp, vp, vp, vp, u32, ctypes.POINTER(f32), vp, vp]),
    "ppf_lpc_adam": (ctypes.c_int, [vp, f32, u32, vp]),
    "ppf_pcg_solve_lpc": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, f32, u32,
                                         ctypes.POINTER(PcgResult), vp]),
    "ppf_pcg_solve": (ctypes.c_int, [vp, vp, vp, vp, vp, f32, u32,
                                     ctypes.POINTER(PcgResult), vp]),
    "ppf_pcg_solve_hist": (ctypes.c_int, [vp, vp, vp, vp, vp, f32, u32, vp,
                                          ctypes.POINTER(PcgResult), vp]),
    "ppf_pcg_spmv": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
    "ppf_pcg_time_spmv": (ctypes.c_int, [vp, vp, vp, vp, vp, u32,
                                         ctypes.POINTER(f32), vp]),
    "ppf_pcg_plan_set_spmv_form": (ctypes.c_int, [vp, ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_int), vp]),
    "ppf_pcg_solve_noalloc": (ctypes.c_int, [u32, u32, vp, vp, vp, vp, vp, vp, f32, u32,
                                             ctypes.POINTER(PcgResult), vp]),
    "ppf_bj_invert": (ctypes.c_int, [u32, vp, vp, vp, vp]),
    "ppf_symm3x3": (ctypes.c_int, [u32, vp, vp, vp, vp]),
    "ppf_narrow_create": (ctypes.c_int, [ctypes.POINTER(vp)]),
    "ppf_narrow_destroy": (None, [vp]),
    "ppf_narrow_build": (ctypes.c_int, [vp, vp, vp]),
    "ppf_narrow_tree": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, ctypes.POINTER(u32),
                                       ctypes.POINTER(u32), vp]),
    "ppf_narrow_proximity": (ctypes.c_int, [vp, vp, vp, ctypes.c_uint64,
                                            ctypes.POINTER(ctypes.c_uint64), vp]),
    "ppf_contact_create": (ctypes.c_int, [ctypes.POINTER(vp)]),
    "ppf_contact_destroy": (None, [vp]),
    "ppf_contact_assemble": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_uint64, vp,
                                            ctypes.POINTER(ctypes.c_uint64), vp]),
    "ppf_contact_blocks": (ctypes.c_int, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                          ctypes.POINTER(vp), ctypes.POINTER(vp),
                                          ctypes.POINTER(ctypes.c_uint64)]),
    "ppf_narrow_static": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_uint64), vp]),
    "ppf_intersection_check": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, u32, vp, u32,
                                              ctypes.POINTER(u32), vp]),
    "ppf_elastic_create": (ctypes.c_int, [vp, vp, ctypes.POINTER(vp)]),
    "ppf_elastic_destroy": (None, [vp]),
    "ppf_elastic_pattern": (ctypes.c_int, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                           ctypes.POINTER(u32)]),
    "ppf_elastic_assemble": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "ppf_elastic_plasticity": (ctypes.c_int, [vp, vp, ctypes.c_float, vp, vp, vp]),
    "ppf_elastic_strain_limit": (ctypes.c_int, [vp, vp, ctypes.c_float, vp, vp, vp]),
    "ppf_elastic_strain_limit_toi": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_float,
                                                    ctypes.POINTER(ctypes.c_float), vp]),
    "ppf_stepper_create": (ctypes.c_int, [vp, vp, ctypes.POINTER(vp)]),
    "ppf_stepper_destroy": (None, [vp]),
    "ppf_stepper_advance": (ctypes.c_int, [vp, vp, vp, vp, vp]),
    "ppf_stepper_advance_scenes": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
    "ppf_stepper_update_pins": (ctypes.c_int, [vp, vp, vp]),
    "ppf_stepper_set_preconditioner": (ctypes.c_int, [vp, vp]),
    "ppf_stepper_system": (ctypes.c_int, [vp, ctypes.POINTER(u32), ctypes.POINTER(u32), vp,
                                          vp, vp, vp, vp, vp]),
    "ppf_stepper_intersection_records": (u32, [vp, vp, u32]),
    "ppf_stepper_contact_pattern": (ctypes.c_int, [vp, ctypes.POINTER(u32), vp, vp, vp, vp]),
    "ppf_ccd_create": (ctypes.c_int, [ctypes.POINTER(vp)]),
    "ppf_ccd_destroy": (None, [vp]),
    "ppf_line_search": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_int,
                                       ctypes.POINTER(f32), ctypes.POINTER(ctypes.c_int),
                                       vp]),
    "ppf_ccd_batch": (ctypes.c_int, [ctypes.c_int, u32, vp, vp, vp, vp, f32, f32, vp,
                                     vp, vp]),
}

_lib = None


def load():
    """Load the HIP library (raises OSError if it is absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it first "
                          "(python -c 'import __graft_entry__ as g; g.build()')")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class PPFError(RuntimeError):
    pass


def check(rc, what):
    if rc != 0:
        msg = load().ppf_last_error().decode(errors="replace")
        raise PPFError(f"{what} failed (code {rc}): {msg}")


def walk_stats_enable(on):
    """Start (zeroing the counters) or stop counting the BVH walks."""
    check(load().ppf_walk_stats_enable(1 if on else 0), "ppf_walk_stats_enable")


def walk_stats():
    """Walk statistics per kind (WALK_KINDS): dicts of launches, queries,
    nodes, leaves, sweeps and ms."""
    out = (WalkStats * 4)()
    check(load().ppf_walk_stats_read(out), "ppf_walk_stats_read")
    return {k: {f: getattr(out[i], f) for f, _ in WalkStats._fields_}
            for i, k in enumerate(WALK_KINDS)}


def alloc_counts():
    """(allocations, frees) of device and pinned-host memory the library has
    made since it was loaded."""
    a, f = ctypes.c_uint64(), ctypes.c_uint64()
    load().ppf_alloc_counts(ctypes.byref(a), ctypes.byref(f))
    return a.value, f.value
