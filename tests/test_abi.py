"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: this runs without a GPU)."""
import glob
import os
import re
import ctypes

from ppf_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in [os.path.join(ROOT, "include", "ppf_mi355x.h")]:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ppf_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert "ppf_pcg_solve" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    names = declared_functions()
    assert names <= set(_lib._PROTOS), names - set(_lib._PROTOS)
    _lib.load()  # prototypes bind


def test_error_string_available_without_gpu():
    lib = _lib.load()
    assert isinstance(lib.ppf_last_error(), bytes)


def _mirrors():
    from ppf_amd import _lib as lib_mod, ccd, elastic, narrow, stepper
    return {"ppf_pcg_result": lib_mod.PcgResult, "ppf_static_mesh": ccd.StaticMeshC,
            "ppf_ls_params": ccd.LsParamsC, "ppf_shell": elastic.ShellC,
            "ppf_elastic_params": elastic.ElasticParamsC,
            "ppf_contact_mesh": narrow.ContactMeshC,
            "ppf_contact_params": narrow.ContactParamsC, "ppf_sphere": stepper.SphereC,
            "ppf_floor": stepper.FloorC, "ppf_scene": stepper.SceneC,
            "ppf_step_params": stepper.StepParamsC, "ppf_step_result": stepper.StepResultC,
            "ppf_walk_stats": lib_mod.WalkStats}


def _c_fields(text, name):
    """Field names of `typedef struct name { ... } name;` in declaration order."""
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        decl = re.sub(r"\[[^\]]*\]", "", decl)
        head, _, rest = decl.partition(",")
        first = re.findall(r"(\w+)\s*$", head)[0]
        fields.append(first)
        fields += [re.sub(r"[\s\*]", "", r) for r in rest.split(",") if r.strip()]
    return fields


def test_ctypes_mirrors_match_c_layout(tmp_path):
    """Every ctypes mirror of a header struct has the C field names, offsets,
    size and alignment (a mirror short of a field hands the library garbage
    in that field's place)."""
    import subprocess
    header = os.path.join(ROOT, "include", "ppf_mi355x.h")
    text = re.sub(r"/\*.*?\*/", "", open(header).read(), flags=re.S)
    mirrors = _mirrors()
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "ppf_mi355x.h"',
             "int main(void) {"]
    for name in sorted(mirrors):
        lines.append(f'printf("{name} size %zu align %zu\\n", sizeof({name}), '
                     f'_Alignof({name}));')
        for f in _c_fields(text, name):
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        parts = line.split()
        got[parts[0]] = (int(parts[2]), int(parts[4])) if parts[1] == "size" else int(parts[1])
    for name, cls in mirrors.items():
        cf = _c_fields(text, name)
        pf = [f[0] for f in cls._fields_]
        assert pf == cf, (name, pf, cf)
        for f in cf:
            assert getattr(cls, f).offset == got[f"{name}.{f}"], f"{name}.{f}"
        assert (ctypes.sizeof(cls), ctypes.alignment(cls)) == got[name], name
