"""The solver-backend boundary (include/ppf_backend.h, libsimbackend_mi355x.so)
without a GPU:

* every #[repr(C)] struct the Rust host passes has, in the C header, the
  size, alignment and field offsets that the C layout rules give the
  reference's data.rs / cvec.rs / cvecvec.rs (tests/golden/abi_layout.json,
  written by tests/golden/make_abi_layout.py from those files);
* the library exports the 19 functions the host imports and the two fatal
  globals, and nothing the host would not find under those names is missing;
* calls that need no device (fatal code / detail, set_log_path,
  fetch_intersection_records before any step) behave as the reference's.
"""
import ctypes
import json
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ppf_backend.h")
LIB = os.path.join(ROOT, "ppf-contact-solver_amd", "libsimbackend_mi355x.so")
LAYOUT = os.path.join(ROOT, "tests", "golden", "abi_layout.json")

# backend.rs:28-62, main.rs:43-45, status_writer.rs:39-46
HOST_IMPORTS = ["advance", "fetch", "fetch_inv_rest", "fetch_rest_angles", "fetch_dyn_counts",
                "fetch_dyn", "update_dyn", "update_constraint", "update_rest_shape",
                "initialize", "override_velocity", "gather_current_positions",
                "override_angular_velocity", "init_collision_windows",
                "refresh_collision_active", "fetch_intersection_records", "set_log_path",
                "ppf_fatal_code", "ppf_fatal_detail"]


def test_layout_matches_host_structs(tmp_path):
    layout = json.load(open(LAYOUT))["structs"]
    lines = ['#include <cstdio>', '#include <cstddef>', '#include "ppf_backend.h"',
             "int main() {"]
    for name, d in sorted(layout.items()):
        c = "cts_" + name
        lines.append(f'printf("{name} size %zu align %zu\\n", sizeof({c}), alignof({c}));')
        for f in d["fields"]:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({c}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.cpp"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        parts = line.split()
        if parts[1] == "size":
            got[parts[0]] = (int(parts[2]), int(parts[4]))
        else:
            got[parts[0]] = int(parts[1])
    assert len(layout) == 40
    for name, d in layout.items():
        assert got[name] == (d["size"], d["align"]), name
        for f, off in d["fields"].items():
            assert got[f"{name}.{f}"] == off, f"{name}.{f}"
    assert layout["DataSet"]["size"] == got["DataSet"][0]


def test_header_declares_exactly_the_host_imports():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    declared = set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(\w+)\s*\([^;{]*\)\s*;", text, re.M))
    declared -= {"CTS_VEC", "CTS_VECVEC"}
    assert declared == set(HOST_IMPORTS), declared ^ set(HOST_IMPORTS)


def test_library_exports_host_imports():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} missing: build first")
    lib = ctypes.CDLL(LIB)
    for name in HOST_IMPORTS + ["g_ppf_fatal_code", "g_ppf_fatal_detail"]:
        assert hasattr(lib, name), name


def test_calls_that_need_no_device():
    lib = ctypes.CDLL(LIB)
    lib.ppf_fatal_code.restype = ctypes.c_ubyte
    lib.ppf_fatal_detail.restype = ctypes.c_char_p
    assert lib.ppf_fatal_code() == 0
    assert lib.ppf_fatal_detail() == b""
    lib.set_log_path(b"/tmp")
    lib.fetch_intersection_records.restype = ctypes.c_uint32
    assert lib.fetch_intersection_records(None, 256) == 0
    # a null DataSet is refused without touching a device, with a detail
    lib.initialize.restype = ctypes.c_bool
    assert lib.initialize(None, None) is False
    assert lib.ppf_fatal_code() == 4
    assert b"null" in lib.ppf_fatal_detail()
