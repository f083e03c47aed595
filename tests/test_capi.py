"""The C-ABI library loads, exports every entry point include/bm3d_mi355x.h
declares, and validates arguments with the reference's messages
(source/source.cpp:503-611). No GPU work: validation precedes any device
call."""
import ctypes
import os
import re

import pytest

from bm3d_amd import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "bm3d_mi355x.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bm3d(?:v2)?_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header():
    lib = C.lib()
    names = header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), f"{n} declared but not exported"
    assert set(C.EXPORTS) <= set(names)


def test_version_and_error_string():
    assert b"gfx950" in C.lib().bm3d_version()
    assert C.lib().bm3d_last_error() is not None


def _desc(**kw):
    d = C.Desc()
    base = dict(width=64, height=48, stride=64, block_step=8, bm_range=9, radius=0, ps_num=2,
                ps_range=4, chroma=0, final_=0, extractor_exp=0, precision=0)
    base.update(kw)
    sig = base.pop("sigma", (3.0, 3.0, 3.0))
    for k, v in base.items():
        setattr(d, k, v)
    for i in range(3):
        d.sigma[i] = sig[i]
    return d


@pytest.mark.parametrize("kw,msg", [
    (dict(sigma=(-1.0, 3.0, 3.0)), 'BM3D: "sigma" must be non-negative'),
    (dict(block_step=0), 'BM3D: "block_step" must be in range [1, 8]'),
    (dict(block_step=9), 'BM3D: "block_step" must be in range [1, 8]'),
    (dict(bm_range=0), 'BM3D: "bm_range" must be positive'),
    (dict(radius=-1), 'BM3D: "radius" must be non-negative'),
    (dict(ps_num=0), 'BM3D: "ps_num" must be in range [1, 8]'),
    (dict(ps_num=9), 'BM3D: "ps_num" must be in range [1, 8]'),
    (dict(ps_range=0), 'BM3D: "ps_range" must be positive'),
    (dict(width=7), "BM3D: plane must be at least 8x8"),
    (dict(stride=10), "BM3D: stride must be >= width"),
])
def test_plan_create_rejects(kw, msg):
    h = ctypes.c_void_p()
    rc = C.lib().bm3d_plan_create(ctypes.byref(_desc(**kw)), ctypes.byref(h))
    assert rc == C.BM3D_EINVAL
    assert C.lib().bm3d_last_error().decode() == msg
    assert not h.value


def test_execute_null_plan():
    rc = C.lib().bm3d_execute(None, None, 0, 0, None, None, 0, 1, None)
    assert rc == C.BM3D_EINVAL
    assert "null" in C.lib().bm3d_last_error().decode()
