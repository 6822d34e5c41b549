"""Single precision on the device: no gfx950 instruction of the built library
computes, compares or converts in fp64.

The reference guards the same property by reading the machine code of the
device image it has just built and failing on any double-precision
instruction, conversions included (crates/ppf-cts-solver/build.rs:223-466):
a double can reach the device without appearing in the source, e.g. through
the slow-path argument reduction of a library trig function. Here the gfx950
code objects of libppf_mi355x.so and libsimbackend_mi355x.so (this repo's own
build) are disassembled with the ROCm LLVM objdump and every instruction
whose mnemonic names f64 (v_add_f64, v_fma_f64, v_cvt_f64_f32,
v_cvt_f32_f64, v_cmp_*_f64, v_rcp_f64, v_div_*_f64, v_trig_preop_f64, ...)
is reported with its kernel. The list of allowed sites is empty. A tiny
kernel with one double multiply, compiled the same way, shows the scan sees
fp64 when it is there.
"""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ppf-contact-solver_amd")
OBJDUMP = "/opt/rocm/llvm/bin/llvm-objdump"
HIPCC = "/opt/rocm/bin/hipcc"
KNOWN_FP64_SITES = []  # kernels allowed to hold fp64: none

_F64 = re.compile(r"f64")
_SYM = re.compile(r"^[0-9a-f]+ <([^>]+)>:")


def _fp64_sites(library, min_instr=1000):
    """{kernel symbol: [fp64 instruction lines]} over the library's gfx950 code
    objects (at least `min_instr` instructions must be seen)."""
    out = {}
    with tempfile.TemporaryDirectory() as d:
        lib = os.path.join(d, os.path.basename(library))
        shutil.copy(library, lib)
        subprocess.run([OBJDUMP, "--offloading", lib], cwd=d, check=True,
                       capture_output=True)
        objs = [os.path.join(d, f) for f in sorted(os.listdir(d)) if f.endswith("gfx950")]
        assert objs, f"no gfx950 code object in {library}"
        ninstr = 0
        for co in objs:
            dis = subprocess.run([OBJDUMP, "-d", "--mcpu=gfx950", co], check=True,
                                 capture_output=True, text=True).stdout
            sym = None
            for line in dis.splitlines():
                m = _SYM.match(line)
                if m:
                    sym = m.group(1)
                    continue
                s = line.strip()
                if not s or s.startswith(";") or sym is None:
                    continue
                mnem = s.split()[0]
                if not re.match(r"^[sv]_|^global_|^buffer_|^ds_|^flat_", mnem):
                    continue
                ninstr += 1
                if _F64.search(mnem):
                    out.setdefault(sym, []).append(s)
        assert ninstr >= min_instr, f"disassembly of {library} looks empty ({ninstr} instructions)"
    return out


@pytest.mark.parametrize("name,floor", [("libppf_mi355x.so", 10000),
                                        ("libsimbackend_mi355x.so", 50)])
def test_no_fp64_in_device_code(name, floor):
    library = os.path.join(PKG, name)
    assert os.path.exists(library), f"{name} is not built (run __graft_entry__.build())"
    sites = {k: v for k, v in _fp64_sites(library, floor).items() if k not in KNOWN_FP64_SITES}
    report = "\n".join(f"{k}: {len(v)} fp64 instructions, e.g. {v[0]}" for k, v in sites.items())
    assert not sites, "fp64 on the device:\n" + report


def test_scan_sees_fp64():
    src = r"""
#include <hip/hip_runtime.h>
extern "C" __global__ void k_planted(const float *a, float *b, double s) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    b[i] = (float)((double)a[i] * s);
}
"""
    with tempfile.TemporaryDirectory() as d:
        hip = os.path.join(d, "planted.hip")
        so = os.path.join(d, "libplanted.so")
        open(hip, "w").write(src)
        subprocess.run([HIPCC, "-O3", "--offload-arch=gfx950", "-fPIC", "-shared", hip, "-o", so],
                       check=True, capture_output=True)
        sites = _fp64_sites(so, min_instr=4)
    assert "k_planted" in sites and any("f64" in s for s in sites["k_planted"])

