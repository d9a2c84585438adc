"""The C ABI: every function include/*.h declares is exported by
libfdtd_amd.so with C linkage, the package's export list matches the
headers, and the shared structs keep the reference's layout."""
import ctypes as C
import os
import re
import subprocess

import pytest

import fdtd_amd as F

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
DECL = re.compile(r"^(?:const\s+)?(?:int|void|char|FdtdParams|FdtdField)\s*\*?\s*(\w+)\s*\(", re.M)


def declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(DECL.findall(text)))


@pytest.mark.parametrize("header", ["fdtd.h", "fdtd_calculations.h", "fdtd_amd.h"])
def test_header_symbols_exported(header):
    names = declared(header)
    assert names, header
    assert sorted(F.EXPORTS[header]) == names
    out = subprocess.run(["nm", "-D", "--defined-only", F.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [n for n in names if n not in exported]
    assert not missing, f"{header}: not exported with C linkage: {missing}"
    L = F.lib()
    for n in names:
        getattr(L, n)


def test_reference_struct_layout():
    # fdtd.h:7-34 and :37-54 on LP64 (offsets as gcc lays out the header)
    assert C.sizeof(F.FdtdParams) == 136
    assert F.FdtdParams.sources.offset == 56 and F.FdtdParams.jz.offset == 88
    assert F.FdtdParams.eps0.offset == 132
    assert C.sizeof(F.FdtdField) == 31 * 8
    assert F.FdtdField.hx.offset == 9 * 8 and F.FdtdField.epsI.offset == 21 * 8 and F.FdtdField.rpzEnd.offset == 30 * 8


def test_no_gpu_fails_loudly():
    """Without a GPU the fused path raises instead of falling back."""
    from fdtd_amd import hip
    if hip.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(F.FdtdError):
        F.Context(F.make_params((8, 8, 8), 1e-12, 1e-3, mu0=1.2566371e-06, eps0=8.841941e-12))


EXT_STRUCTS = {
    "FdtdExt": [],
    "FdtdInstance": ["rp", "nsrc", "jz", "steps", "probes"],
    "FdtdBatchInfo": ["step", "tile_y", "grid", "cells"],
    "FdtdServeConfig": ["max_slots", "kchunk"],
    "FdtdRequestInfo": ["t_submit", "cells", "steps"],
    "FdtdServeStats": ["t_first_submit", "segments", "peak_bytes", "peak_running"],
}


def test_extension_struct_layout(tmp_path):
    """The ctypes mirrors of fdtd_amd.h's structs match what a C compiler
    lays out (sizes and field offsets)."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fdtd_amd.h"', "int main(void) {"]
    for name, fields in EXT_STRUCTS.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for f in fields:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}", "-I/opt/rocm/include", str(src),
                    "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True)
               .stdout.splitlines())
    for name, fields in EXT_STRUCTS.items():
        S = getattr(F, name)
        assert int(got[name]) == C.sizeof(S), name
        for f in fields:
            assert int(got[f"{name}.{f}"]) == getattr(S, f).offset, f"{name}.{f}"


def test_constants_match_header():
    """Every FDTD_* value the package mirrors equals the header's #define."""
    text = open(os.path.join(INCLUDE, "fdtd_amd.h")).read()
    defs = {m.group(1): int(m.group(2)) for m in re.finditer(r"^#define FDTD_(\w+)\s+\(?(-?\d+)\)?", text, re.M)}
    mirrored = [n for n in defs if hasattr(F, n)]
    assert len(mirrored) >= 19, mirrored
    for n in mirrored:
        assert getattr(F, n) == defs[n], n
    assert [defs[f"{f.upper()}"] for f in F.FIELDS] == list(range(9))
