"""Writes abi_layout.json: the byte layout (size, alignment, field offsets)
of every #[repr(C)] struct the Rust host passes through the solver FFI, as
the C layout rules place them. Field order and types are read as text from
the reference's crates/ppf-cts-solver/src/{data.rs,cvec.rs,cvecvec.rs}; the
nalgebra aliases at data.rs:11-18 are plain column-major f32/u32 arrays.

Run in a container that holds the reference checkout:
    python tests/golden/make_abi_layout.py /root/reference
"""
import json
import os
import re
import sys

PRIM = {"f32": (4, 4), "f64": (8, 8), "u32": (4, 4), "u8": (1, 1), "bool": (1, 1)}
NALGEBRA = {"Vec3f": (12, 4), "Vec2u": (8, 4), "Vec3u": (12, 4), "Vec4u": (16, 4),
            "Vec6u": (24, 4), "Vec6f": (24, 4), "Mat2x2f": (16, 4), "Mat3x3f": (36, 4)}
PTR = (8, 8)


def parse(src):
    structs, enums = {}, set()
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[[^\]]*\]\s*)*pub (struct|enum) (\w+)(?:<T>)?\s*\{(.*?)\n\}",
                         src, re.S):
        kind, name, body = m.groups()
        if kind == "enum":
            enums.add(name)
            continue
        fields = re.findall(r"^\s*pub (\w+):\s*([^,\n]+),", body, re.M)
        structs[name] = [(f, t.strip()) for f, t in fields]
    return structs, enums


def layout(structs, enums):
    out = {}

    def size_align(t):
        if t in PRIM:
            return PRIM[t]
        if t in NALGEBRA:
            return NALGEBRA[t]
        if t in enums:
            return (4, 4)
        a = re.fullmatch(r"\[(\w+);\s*(\d+)\]", t)
        if a:
            s, al = size_align(a.group(1))
            return (s * int(a.group(2)), al)
        if re.fullmatch(r"\*mut \w+", t):
            return PTR
        g = re.fullmatch(r"(CVec|CVecVec)<.+>", t)
        if g:
            return size_align(g.group(1))
        if t in structs:
            d = struct(t)
            return (d["size"], d["align"])
        raise KeyError(t)

    def struct(name):
        if name in out:
            return out[name]
        off, al, fields = 0, 1, {}
        for f, t in structs[name]:
            s, a = size_align(t)
            off = (off + a - 1) // a * a
            fields[f] = off
            off += s
            al = max(al, a)
        size = (off + al - 1) // al * al
        out[name] = {"size": size, "align": al, "fields": fields}
        return out[name]

    for n in structs:
        struct(n)
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "crates", "ppf-cts-solver", "src")
    text = "".join(open(os.path.join(src, f)).read() for f in ("cvec.rs", "cvecvec.rs", "data.rs"))
    structs, enums = parse(text)
    for g in ("CVec", "CVecVec"):
        structs[g] = [(f, "*mut T" if t.startswith("*mut") else t) for f, t in structs[g]]
    res = layout(structs, enums)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abi_layout.json")
    with open(path, "w") as fh:
        json.dump({"source": "crates/ppf-cts-solver/src/data.rs, cvec.rs, cvecvec.rs",
                   "enums": sorted(enums), "structs": res}, fh, indent=1, sort_keys=True)
    print(f"wrote {path}: {len(res)} structs")


if __name__ == "__main__":
    main()
