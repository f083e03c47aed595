"""The ctypes mirrors of the C-ABI structs (bm3d_amd/_capi.py) have the
layout the header declares: every field at the offset the C compiler gives
it in include/bm3d_mi355x.h, and every struct of the same size. A field
added on one side only (or in a different place) fails here, on the CPU,
instead of corrupting arguments on the GPU."""
import ctypes
import os
import shutil
import subprocess

import pytest

from bm3d_amd import _capi as C

ROOT = os.path.join(os.path.dirname(__file__), os.pardir)
STRUCTS = {
    "bm3d_desc": C.Desc,
    "bm3d_clip": C.Clip,
    "bm3d_video_info": C.VideoInfo,
    "bm3d_args": C.Args,
    "bm3d_frames": C.Frames,
    "bm3d_host_frame": C.HostFrame,
    "bm3d_tuning": C.Tuning,
    "bm3d_batch_job": C.BatchJob,
}


def _c_layout(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bm3d_mi355x.h"', "int main(void) {"]
    for name, cls in STRUCTS.items():
        lines.append(f'    printf("{name} sizeof %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'    printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    layout = {}
    for line in out.splitlines():
        struct, field, value = line.split()
        layout[(struct, field)] = int(value)
    return layout


def test_ctypes_structs_match_the_header(tmp_path):
    layout = _c_layout(tmp_path)
    for name, cls in STRUCTS.items():
        assert ctypes.sizeof(cls) == layout[(name, "sizeof")], name
        for field, _ in cls._fields_:
            assert getattr(cls, field).offset == layout[(name, field)], (name, field)
