"""The ctypes mirrors of the C ABI's structs have the header's layout: every
size and field offset equals what a C compiler makes of
include/occastft_hip.h (a small program compiled against it prints them)."""
import ctypes
import os
import shutil
import subprocess

import pytest

from occastft_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# C struct name -> (ctypes mirror, {ctypes field: C field} where they differ)
STRUCTS = {
    "ostft_params": (N.Params, {}),
    "ostft_geometry": (N.Geometry, {}),
    "ostft_job": (N.Job, {"in_ptr": "in", "out_ptr": "out"}),
    "ostft_server_config": (N.ServerConfig, {}),
    "ostft_server_stats": (N.ServerStats, {}),
}


def test_struct_layouts_match_the_header(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "occastft_hip.h"', "int main(void) {"]
    for cname, (py, rename) in STRUCTS.items():
        lines.append(f'printf("{cname} . %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {rename.get(f, f)}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    seen = 0
    for cname, field, value in zip(out[0::3], out[1::3], out[2::3]):
        py = STRUCTS[cname][0]
        want = ctypes.sizeof(py) if field == "." else getattr(py, field).offset
        assert int(value) == want, f"{cname}.{field}: C {value}, ctypes {want}"
        seen += 1
    assert seen == sum(1 + len(py._fields_) for py, _ in STRUCTS.values())
