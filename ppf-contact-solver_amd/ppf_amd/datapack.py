"""Scene files for the C++ mock host (host/mock_host.cpp).

A pack is a flat list of named arrays:
    b"PPFPACK1", u32 count, then per array:
    u32 name length, name bytes, u8 kind ('f' float32, 'u' uint32, 'd' float64),
    u64 element count, the elements (little endian).

`scene_arrays` turns a scene dict (the scenes.drape_scene layout: contact mesh,
shell, static collision mesh, per-element friction, pins) plus ParamSet values
and a step count into the arrays the mock host reads to fill the host's
DataSet / ParamSet the way crates/ppf-cts-solver/src/builder.rs would: one
parameter record per element, FaceParam model Baraff-Witkin, props from the
shell (mass, area, rest angle, hinge length and area).
"""
import struct

import numpy as np

MAGIC = b"PPFPACK1"
_KINDS = {np.dtype(np.float32): b"f", np.dtype(np.uint32): b"u", np.dtype(np.float64): b"d"}
_DTYPES = {b"f": np.float32, b"u": np.uint32, b"d": np.float64}


def write_pack(path, arrays):
    with open(path, "wb") as fh:
        fh.write(MAGIC)
        fh.write(struct.pack("<I", len(arrays)))
        for name, a in arrays.items():
            a = np.ascontiguousarray(a)
            if a.dtype not in _KINDS:
                raise TypeError(f"{name}: unsupported dtype {a.dtype}")
            nb = name.encode()
            fh.write(struct.pack("<I", len(nb)) + nb + _KINDS[a.dtype])
            fh.write(struct.pack("<Q", a.size))
            fh.write(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes())


def read_pack(path):
    out = {}
    with open(path, "rb") as fh:
        if fh.read(8) != MAGIC:
            raise ValueError(f"{path}: not a pack")
        (count,) = struct.unpack("<I", fh.read(4))
        for _ in range(count):
            (ln,) = struct.unpack("<I", fh.read(4))
            name = fh.read(ln).decode()
            kind = fh.read(1)
            (n,) = struct.unpack("<Q", fh.read(8))
            dt = np.dtype(_DTYPES[kind])
            out[name] = np.frombuffer(fh.read(n * dt.itemsize), dtype=dt.newbyteorder("<"))
    return out


PARAM_KEYS = ("dt", "prev_dt", "cg_tol", "cg_max_iter", "target_toi", "line_search_max_t",
              "ccd_eps", "max_dx", "friction_eps", "friction_mode", "min_newton_steps",
              "max_newton_steps", "air_density", "air_friction", "disable_contact",
              "constraint_tol")


def scene_arrays(sc, params, steps, fix_schedule=None, delta_schedule=None, windows=None):
    """Arrays for the mock host. fix_schedule / delta_schedule: (steps, pins, 3)
    pin positions and step deltas to send with update_constraint before each
    step (pins = the vertices with mesh vflag bit0); windows: optional
    (vert_dmap (nv,), windows (groups, 8, 2), counts (groups,)) collision
    windows."""
    m, sh = sc["mesh"], sc["shell"]
    f32 = lambda a: np.ascontiguousarray(a, np.float32).reshape(-1)
    u32 = lambda a: np.ascontiguousarray(a, np.uint32).reshape(-1)
    a = {}
    a["x"] = f32(m["x"])
    a["prev"] = f32(sc["prev"])
    for k in ("edge", "face", "v2e_off", "v2e", "v2f_off", "v2f", "e2f_off", "e2f"):
        a[k] = u32(m[k])
    a["hinge"] = u32(sh["hinge"])
    a["v_mass"], a["v_area"] = f32(sh["vmass"]), f32(sh["varea"])
    a["v_ghat"], a["v_off"], a["v_fric"] = f32(m["v_ghat"]), f32(m["v_off"]), f32(sc["fric"]["v"])
    a["v_collider"] = u32((np.asarray(m["vflag"]) >> 1) & 1)
    e = np.asarray(m["edge"], np.int64).reshape(-1, 2)
    x = np.asarray(m["x"], np.float64).reshape(-1, 3)
    a["e_mass"], a["e_length"] = f32(sc["emass"]), f32(np.linalg.norm(x[e[:, 0]] - x[e[:, 1]], axis=1))
    a["e_ghat"], a["e_off"], a["e_fric"] = f32(m["e_ghat"]), f32(m["e_off"]), f32(sc["fric"]["e"])
    a["f_mass"], a["f_mu"], a["f_lambda"] = f32(sh["fmass"]), f32(sh["f_mu"]), f32(sh["f_lambda"])
    a["f_ghat"], a["f_off"], a["f_fric"] = f32(m["f_ghat"]), f32(m["f_off"]), f32(sc["fric"]["f"])
    a["inv_rest"] = f32(sh["inv_rest"])
    for k in ("f_strainlimit", "f_plasticity", "f_plastic_threshold", "h_plasticity",
              "h_plastic_threshold"):
        if sh.get(k) is not None:
            a[k] = f32(sh[k])
    for k in ("h_rest", "h_length", "h_area", "h_bend"):
        a[k] = f32(sh[k])
    pins = np.nonzero(np.asarray(m["vflag"]) & 1)[0].astype(np.uint32)
    a["pin_index"] = pins
    if len(pins):
        a["pin_pos"] = f32(np.asarray(sc["fix_pos"], np.float32).reshape(-1, 3)[pins])
        a["pin_delta"] = f32(np.asarray(sc.get("fix_delta", np.zeros((len(x), 3))),
                                        np.float32).reshape(-1, 3)[pins])
    if fix_schedule is not None:
        a["pin_schedule"] = f32(fix_schedule)
        a["pin_delta_schedule"] = f32(delta_schedule if delta_schedule is not None
                                      else np.zeros_like(fix_schedule))
    if sc.get("spheres"):
        a["spheres"] = f32([[*d["center"], d["ghat"], d["friction"], d["radius"],
                             d.get("thickness", 0.0), float(d.get("bowl", False)),
                             float(d.get("reverse", False)), float(d.get("kinematic", False))]
                            for d in sc["spheres"]])
    if sc.get("floors"):
        a["floors"] = f32([[*d["ground"], d["ghat"], d["friction"], d.get("thickness", 0.0),
                            *d.get("up", (0.0, 1.0, 0.0)), float(d.get("kinematic", False))]
                           for d in sc["floors"]])
    st = sc.get("static")
    if st:
        a["s_x"], a["s_edge"], a["s_face"] = f32(st["x"]), u32(st["edge"]), u32(st["face"])
        for k in ("v", "e", "f"):
            a[f"s_{k}_ghat"], a[f"s_{k}_off"] = f32(st[k + "_ghat"]), f32(st[k + "_off"])
            a[f"s_{k}_fric"] = f32(st[k + "_fric"])
    if windows is not None:
        a["cw_dmap"], a["cw_windows"], a["cw_counts"] = u32(windows[0]), f32(windows[1]), u32(windows[2])
    a["gravity"] = f32(params["gravity"])
    a["wind"] = f32(params["wind"])
    for k in PARAM_KEYS:
        a[k] = np.array([params[k]], np.float64)
    if "eiganalysis_eps" in params:
        a["eiganalysis_eps"] = np.array([params["eiganalysis_eps"]], np.float64)
    a["steps"] = np.array([steps], np.uint32)
    return a
